// mps_device.h -- device functions of the explicit hex8/tet4 neo-Hookean step.
//
// Element work per hex8 (2x2x2 Gauss) uses the trilinear ("butterfly") structure of
// the isoparametric map: the 8 nodal vectors are transformed once per element into
// 7 derivative coefficient vectors (xi, eta, zeta, xi*eta, eta*zeta, xi*zeta,
// xi*eta*zeta), from which the Jacobian and displacement gradient at each Gauss point
// follow with adds only; nodal forces are accumulated as 7 moment vectors and
// transformed back once.  In exact arithmetic this equals the reference sequence
// dphi = natDev*J^-1 (classLinearHexahedra.cpp:301-332), F = I + sum_a u_a (x) dphi_a
// (constitutiveManager.cpp:150-164) and f_a = wJ*P*dphi_a (constitutiveAlgorithms.cpp:46-56);
// only the rounding differs.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "mps_common.h"

#define MPS_GP 0.577350269  // hex Gauss abscissa literal, GPsDistribution.cpp:392-432
#define MPS_TET_W 0.166666666667  // tet weight literal, GPsDistribution.cpp:342

#ifndef MPS_STATUS_CODES
#define MPS_STATUS_CODES
#define MPSD_RUNNING 0
#define MPSD_DONE 1
#define MPSD_ERR_DT (-3)
#endif

namespace mpsd {

__device__ __forceinline__ double bits2d(unsigned long long b) { return __longlong_as_double((long long)b); }
__device__ __forceinline__ unsigned long long d2bits(double d) { return (unsigned long long)__double_as_longlong(d); }

// Node sign table of the reference hex (classLinearHexahedra.cpp:346-353).
__device__ __forceinline__ constexpr int sxi(int a) { return (a == 1 || a == 2 || a == 5 || a == 6) ? 1 : -1; }
__device__ __forceinline__ constexpr int seta(int a) { return (a == 2 || a == 3 || a == 6 || a == 7) ? 1 : -1; }
__device__ __forceinline__ constexpr int szeta(int a) { return (a >= 4) ? 1 : -1; }

// ---------------------------------------------------------------------------
// 3x3 helpers (row-major)
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T det3(const T* A) {
    return A[0] * (A[4] * A[8] - A[7] * A[5]) - A[1] * (A[3] * A[8] - A[6] * A[5]) +
           A[2] * (A[3] * A[7] - A[6] * A[4]);
}

__device__ __forceinline__ double rcp(double x);
__device__ __forceinline__ float rcp(float x);

// Inverse by cofactors; 1/det by the hardware reciprocal and Newton steps (rcp below).
template <typename T>
__device__ __forceinline__ void inv3(const T* A, T* R, T det) {
    T idet = rcp(det);
    R[0] = idet * (A[4] * A[8] - A[5] * A[7]);
    R[3] = -idet * (A[3] * A[8] - A[5] * A[6]);
    R[6] = idet * (A[3] * A[7] - A[4] * A[6]);
    R[1] = -idet * (A[1] * A[8] - A[2] * A[7]);
    R[4] = idet * (A[0] * A[8] - A[2] * A[6]);
    R[7] = -idet * (A[0] * A[7] - A[1] * A[6]);
    R[2] = idet * (A[1] * A[5] - A[2] * A[4]);
    R[5] = -idet * (A[0] * A[5] - A[2] * A[3]);
    R[8] = idet * (A[0] * A[4] - A[1] * A[3]);
}

// Trilinear shape values and natural derivatives nd[a*3 + d] at (xi, eta, zeta)
// (classLinearHexahedra::getInterpolationParameters, classLinearHexahedra.cpp:344-381).
__host__ __device__ inline void hex_shape_ref(double xi, double eta, double zeta, double* phi, double* nd) {
#pragma clang fp contract(off)
    const double x = xi, y = eta, z = zeta;
    phi[0] = 0.125 * (1 - x) * (1 - y) * (1 - z);
    phi[1] = 0.125 * (1 + x) * (1 - y) * (1 - z);
    phi[2] = 0.125 * (1 + x) * (1 + y) * (1 - z);
    phi[3] = 0.125 * (1 - x) * (1 + y) * (1 - z);
    phi[4] = 0.125 * (1 - x) * (1 - y) * (1 + z);
    phi[5] = 0.125 * (1 + x) * (1 - y) * (1 + z);
    phi[6] = 0.125 * (1 + x) * (1 + y) * (1 + z);
    phi[7] = 0.125 * (1 - x) * (1 + y) * (1 + z);
    nd[0] = -0.125 * (1 - y) * (1 - z);
    nd[3] = 0.125 * (1 - y) * (1 - z);
    nd[6] = 0.125 * (1 + y) * (1 - z);
    nd[9] = -0.125 * (1 + y) * (1 - z);
    nd[12] = -0.125 * (1 - y) * (1 + z);
    nd[15] = 0.125 * (1 - y) * (1 + z);
    nd[18] = 0.125 * (1 + y) * (1 + z);
    nd[21] = -0.125 * (1 + y) * (1 + z);
    nd[1] = -0.125 * (1 - x) * (1 - z);
    nd[4] = -0.125 * (1 + x) * (1 - z);
    nd[7] = 0.125 * (1 + x) * (1 - z);
    nd[10] = 0.125 * (1 - x) * (1 - z);
    nd[13] = -0.125 * (1 - x) * (1 + z);
    nd[16] = -0.125 * (1 + x) * (1 + z);
    nd[19] = 0.125 * (1 + x) * (1 + z);
    nd[22] = 0.125 * (1 - x) * (1 + z);
    nd[2] = -0.125 * (1 - x) * (1 - y);
    nd[5] = -0.125 * (1 + x) * (1 - y);
    nd[8] = -0.125 * (1 + x) * (1 + y);
    nd[11] = -0.125 * (1 - x) * (1 + y);
    nd[14] = 0.125 * (1 - x) * (1 - y);
    nd[17] = 0.125 * (1 + x) * (1 - y);
    nd[20] = 0.125 * (1 + x) * (1 + y);
    nd[23] = 0.125 * (1 - x) * (1 + y);
}

// determinantTensor with no FMA contraction (maths.cpp:1564-1577), for the set-up path.
__host__ __device__ inline double det3_strict(const double* A) {
#pragma clang fp contract(off)
    return A[0] * (A[4] * A[8] - A[7] * A[5]) - A[1] * (A[3] * A[8] - A[6] * A[5]) +
           A[2] * (A[3] * A[7] - A[6] * A[4]);
}

// det(I+H) - 1 without cancellation: tr H + i2(H) + det H.
template <typename T>
__device__ __forceinline__ T detIpH_m1(const T* H) {
    T tr = H[0] + H[4] + H[8];
    T i2 = H[0] * H[4] + H[4] * H[8] + H[0] * H[8] - H[1] * H[3] - H[5] * H[7] - H[2] * H[6];
    return tr + i2 + det3(H);
}

// Cofactor matrix cof(A)(i,j) = (-1)^(i+j) minor_ij; A^-1 = cof(A)^T / det A.
template <typename T>
__device__ __forceinline__ void cof3(const T* A, T* C) {
    C[0] = A[4] * A[8] - A[5] * A[7];
    C[1] = A[5] * A[6] - A[3] * A[8];
    C[2] = A[3] * A[7] - A[4] * A[6];
    C[3] = A[2] * A[7] - A[1] * A[8];
    C[4] = A[0] * A[8] - A[2] * A[6];
    C[5] = A[1] * A[6] - A[0] * A[7];
    C[6] = A[1] * A[5] - A[2] * A[4];
    C[7] = A[2] * A[3] - A[0] * A[5];
    C[8] = A[0] * A[4] - A[1] * A[3];
}

// Reciprocal: hardware estimate refined by Newton steps (<= 1 ulp for FP64).
__device__ __forceinline__ double rcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(e, r, r);
    e = fma(-x, r, 1.0);
    return fma(e, r, r);
}
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// log(J) near J = 1 (FP64).  With x = J - 1 (exact for J in [0.5, 2]) and t = x/(2+x),
// log(1+x) = 2 atanh(t) = 2 (t + t^3/3 + ... + t^15/15); for |x| < 0.1, |t| < 0.0527 and the
// truncation error is below 2^-60 relative.  Outside that range the library log is used.
__device__ __forceinline__ double log_near1(double J) {
    const double x = J - 1.0;
    if (fabs(x) >= 0.1) return log(J);
    const double t = x * rcp(2.0 + x);
    const double t2 = t * t;
    double p = 1.0 / 15.0;
    p = fma(p, t2, 1.0 / 13.0);
    p = fma(p, t2, 1.0 / 11.0);
    p = fma(p, t2, 1.0 / 9.0);
    p = fma(p, t2, 1.0 / 7.0);
    p = fma(p, t2, 1.0 / 5.0);
    p = fma(p, t2, 1.0 / 3.0);
    return 2.0 * fma(p * t2, t, t);
}

// log(1+x) (FP32): Taylor series to x^9 for |x| < 0.05 (truncation < 2e-13 relative);
// otherwise Goldberg's form log(u) * x / (u - 1), u = 1 + x.
__device__ __forceinline__ float log1p_fast(float x) {
    if (fabsf(x) < 0.05f) {
        float p = 1.f / 9.f;
        p = fmaf(p, x, -1.f / 8.f);
        p = fmaf(p, x, 1.f / 7.f);
        p = fmaf(p, x, -1.f / 6.f);
        p = fmaf(p, x, 1.f / 5.f);
        p = fmaf(p, x, -1.f / 4.f);
        p = fmaf(p, x, 1.f / 3.f);
        p = fmaf(p, x, -1.f / 2.f);
        return fmaf(p * x, x, x);
    }
    const float u = 1.f + x;
    return __logf(u) * (x / (u - 1.f));
}

// ---------------------------------------------------------------------------
// Constitutive laws: first Piola-Kirchhoff stress P(H), H = F - I.
//  3D compressible NH (hyperElasticNeoHookean.cpp:235-257): S = mu I + (lambda lnJ - mu) C^-1,
//     P = F S.  Since F C^-1 = F^-T this is P = mu F + (lambda lnJ - mu) F^-T, F^-T = cof(F)/J.
//  Belytschko NH (hyperElasticNeoHookean.cpp:88-126): sigma = mu/J^{5/3} B + (-(mu/3) trB/J^{5/3}
//     + K (J-1)) I, P = J sigma F^-T = sigma cof(F).
// ---------------------------------------------------------------------------
//  St-Venant-Kirchhoff (hyperElasticStVenantKirchhoff.cpp:203-226): P = F S, S = lambda tr(E) I + 2 mu E,
//     E = (F^T F - I)/2 = (H + H^T + H^T H)/2.
//  Linear elastic (linearElastic.cpp:63-95): P = lambda tr(e) I + 2 mu e, e = sym(F) - I = (H + H^T)/2.
// Both are written in H, which carries no cancellation at small strain.  (pa, pb) = (lambda, mu).
template <typename T, int LAW>
__device__ __forceinline__ void pk1_svk_lin(const T* H, T lam, T mu, T* P) {
    T E[9];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            T e = T(0.5) * (H[i * 3 + j] + H[j * 3 + i]);
            if (LAW == 2) e += T(0.5) * (H[i] * H[j] + H[3 + i] * H[3 + j] + H[6 + i] * H[6 + j]);
            E[i * 3 + j] = e;
        }
    const T ltr = lam * (E[0] + E[4] + E[8]);
    const T mu2 = T(2) * mu;
    T S[9];
#pragma unroll
    for (int i = 0; i < 9; i++) S[i] = mu2 * E[i];
    S[0] += ltr;
    S[4] += ltr;
    S[8] += ltr;
    if (LAW == 3) {
#pragma unroll
        for (int i = 0; i < 9; i++) P[i] = S[i];
        return;
    }
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            P[i * 3 + j] = S[i * 3 + j] + H[i * 3] * S[j] + H[i * 3 + 1] * S[3 + j] + H[i * 3 + 2] * S[6 + j];
}

__device__ __forceinline__ double defo_energy(int law, const double* F, double lambda, double mu);

// FP64: F-based, as the reference (neo-Hookean laws).  With W and Jout (energy steps) also the energy density
// W(F) (defo_energy) and J = det F, from the quantities P is computed with; P is the same either way.
template <int LAW>
__device__ __forceinline__ void pk1(const double* H, double pa, double pb, double* P, double* W = nullptr,
                                    double* Jout = nullptr) {
    if (LAW >= 2) {
        pk1_svk_lin<double, LAW>(H, pa, pb, P);
        if (W) {
            double F[9];
#pragma unroll
            for (int i = 0; i < 9; i++) F[i] = H[i] + ((i % 4) == 0 ? 1.0 : 0.0);
            *W = defo_energy(LAW, F, pa, pb);
            *Jout = 1.0 + detIpH_m1(H);
        }
        return;
    }
    double F[9];
#pragma unroll
    for (int i = 0; i < 9; i++) F[i] = H[i];
    F[0] += 1.0;
    F[4] += 1.0;
    F[8] += 1.0;
    double cF[9];
    cof3(F, cF);
    const double J = F[0] * cF[0] + F[1] * cF[1] + F[2] * cF[2];
    if (Jout) *Jout = J;
    if (LAW == 0) {  // pa = lambda, pb = mu
        const double l = log_near1(J);
        const double alpha = pa * l - pb;
        const double s = alpha * rcp(J);
#pragma unroll
        for (int i = 0; i < 9; i++) P[i] = pb * F[i] + s * cF[i];
        if (W) {  // 1/2 lambda (ln J)^2 - mu ln J + 1/2 mu tr C (hyperElasticNeoHookean.cpp:291-304)
            double trC = 0.0;
#pragma unroll
            for (int i = 0; i < 9; i++) trC += F[i] * F[i];
            *W = 0.5 * pa * l * l - pb * l + 0.5 * pb * trC;
        }
    } else {  // pa = mu, pb = K1 (no defoEnergy: W = 0, constitutiveModels.h:91-94)
        if (W) *W = 0.0;
        double B[9];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = i; j < 3; j++) {
                B[i * 3 + j] = F[i * 3] * F[j * 3] + F[i * 3 + 1] * F[j * 3 + 1] + F[i * 3 + 2] * F[j * 3 + 2];
                B[j * 3 + i] = B[i * 3 + j];
            }
        const double trB = B[0] + B[4] + B[8];
        // J^(-5/3) = exp(-5/3 ln J): a few ulp from 1/pow(J, 5/3), with far fewer registers than the
        // inlined pow (which held the FP64 Belytschko kernel at one wave per SIMD)
        const double rf1 = exp(-(5.0 / 3.0) * log_near1(J));
        const double f2 = pa * rf1;
        const double f3 = -(pa / 3.0) * trB * rf1 + pb * (J - 1.0);
        double sg[9];
#pragma unroll
        for (int i = 0; i < 9; i++) sg[i] = f2 * B[i];
        sg[0] += f3;
        sg[4] += f3;
        sg[8] += f3;
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++)
                P[i * 3 + j] = sg[i * 3] * cF[j] + sg[i * 3 + 1] * cF[3 + j] + sg[i * 3 + 2] * cF[6 + j];
    }
}

// FP32: H-based so that small strains carry no cancellation:
//   J - 1 = tr H + i2(H) + det H,  B - I = H + H^T + H H^T.
// Compressible: cof(F) = (1 + tr H) I - H^T + cof(H), so with K = H^T - cof(H) and s = i2(H) + det H
// (i2 = tr cof(H), det H = H(0,:).cof(H)(0,:)), F^-T = ((1 + tr H) I - K)/J, F - F^-T = H + (K + s I)/J
// and P = mu (F - F^-T) + lambda lnJ F^-T = mu H + a K + b I, a = (mu - lambda lnJ)/J,
// b = (mu s + lambda lnJ (1 + tr H))/J: every term is of the order of the strain.
template <int LAW>
__device__ __forceinline__ void pk1(const float* H, float pa, float pb, float* P) {
    if (LAW >= 2) {
        pk1_svk_lin<float, LAW>(H, pa, pb, P);
        return;
    }
    if (LAW == 0) {  // pa = lambda, pb = mu
        float cH[9];
        cof3(H, cH);
        const float tr = H[0] + H[4] + H[8];
        const float sv = (cH[0] + cH[4] + cH[8]) + (H[0] * cH[0] + H[1] * cH[1] + H[2] * cH[2]);
        const float Jm1 = tr + sv;
        const float lt = pa * log1p_fast(Jm1);
        const float rJ = rcp(1.f + Jm1);
        const float a = (pb - lt) * rJ;
        const float b = (pb * sv + lt * (1.f + tr)) * rJ;
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) P[i * 3 + j] = pb * H[i * 3 + j] + a * (H[j * 3 + i] - cH[i * 3 + j]);
        P[0] += b;
        P[4] += b;
        P[8] += b;
        return;
    }
    const float Jm1 = detIpH_m1(H);
    const float lnJ = log1p_fast(Jm1);
    float F[9];
#pragma unroll
    for (int i = 0; i < 9; i++) F[i] = H[i];
    F[0] += 1.f;
    F[4] += 1.f;
    F[8] += 1.f;
    float cF[9];
    cof3(F, cF);
    {  // Belytschko (pa = mu, pb = K1): sigma = mu J^{-5/3} dev(B - I) + K (J-1) I ; P = sigma cof(F)
        float Bm[9];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = i; j < 3; j++) {
                Bm[i * 3 + j] = H[i * 3 + j] + H[j * 3 + i] + H[i * 3] * H[j * 3] + H[i * 3 + 1] * H[j * 3 + 1] +
                                H[i * 3 + 2] * H[j * 3 + 2];
                Bm[j * 3 + i] = Bm[i * 3 + j];
            }
        const float tr3 = (Bm[0] + Bm[4] + Bm[8]) * (1.f / 3.f);
        const float f2 = pa * __expf(-(5.f / 3.f) * lnJ);
        const float vol = pb * Jm1 - f2 * tr3;
        float sg[9];
#pragma unroll
        for (int i = 0; i < 9; i++) sg[i] = f2 * Bm[i];
        sg[0] += vol;
        sg[4] += vol;
        sg[8] += vol;
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++)
                P[i * 3 + j] = sg[i * 3] * cF[j] + sg[i * 3 + 1] * cF[3 + j] + sg[i * 3 + 2] * cF[6 + j];
    }
}

// Deformation energy W(F): hyperElasticNeoHookean.cpp:291-304 (no -3mu/2 constant);
// the Belytschko law has no defoEnergy override and contributes 0 (constitutiveModels.h:91-94);
// St-Venant-Kirchhoff 0.5 S:E (hyperElasticStVenantKirchhoff.cpp:251-274); linear elastic
// 0.5 (P:F - P:I) (linearElastic.cpp:97-110).
__device__ __forceinline__ double defo_energy(int law, const double* F, double lambda, double mu) {
    if (law == 1) return 0.0;
    if (law >= 2) {
        double H[9], P[9];
#pragma unroll
        for (int i = 0; i < 9; i++) H[i] = F[i] - ((i % 4) == 0 ? 1.0 : 0.0);
        double w = 0.0;
        if (law == 2) {
            double E[9], S[9];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++)
                    E[i * 3 + j] = 0.5 * (H[i * 3 + j] + H[j * 3 + i] + H[i] * H[j] + H[3 + i] * H[3 + j] + H[6 + i] * H[6 + j]);
            const double ltr = lambda * (E[0] + E[4] + E[8]);
            for (int i = 0; i < 9; i++) S[i] = 2.0 * mu * E[i] + ((i % 4) == 0 ? ltr : 0.0);
            for (int i = 0; i < 9; i++) w += S[i] * E[i];
            return 0.5 * w;
        }
        pk1_svk_lin<double, 3>(H, lambda, mu, P);
        for (int i = 0; i < 9; i++) w += P[i] * H[i];  // P:F - P:I = P:H
        return 0.5 * w;
    }
    double J = det3(F);
    double trC = 0;
#pragma unroll
    for (int i = 0; i < 3; i++) trC += F[i] * F[i] + F[3 + i] * F[3 + i] + F[6 + i] * F[6 + i];
    double l = log_near1(J);
    return 0.5 * lambda * l * l - mu * l + 0.5 * mu * trC;
}

// ---------------------------------------------------------------------------
// Characteristic length (always FP64, no FMA contraction: same rounding as the
// reference geometry.cpp, so dt is bitwise reproducible for identical u).
// ---------------------------------------------------------------------------
// areaTriangle 3D: geometry.cpp:112-130 ; p, q, r are vertices 1, 2, 3.  tri_norm2 is the squared norm
// of the cross product; the area is 0.5 * its square root.
__device__ __forceinline__ double tri_norm2(const double* p, const double* q, const double* r) {
#pragma clang fp contract(off)
    double A0 = p[0] - r[0], A1 = p[1] - r[1], A2 = p[2] - r[2];
    double B0 = q[0] - r[0], B1 = q[1] - r[1], B2 = q[2] - r[2];
    double C0 = A1 * B2 - A2 * B1;
    double C1 = A2 * B0 - A0 * B2;
    double C2 = A0 * B1 - A1 * B0;
    return C0 * C0 + C1 * C1 + C2 * C2;
}

// NS square roots in place, bitwise those of sqrt().  sqrt() on this target is correctly rounded: the
// hardware reciprocal square root and two Newton corrections, with a scaling by 2^256 below 2^-767 and a
// final test for zero and infinity.  When every argument of every active lane of the wave lies in
// [2^-767, inf), the scaling is the identity and the test passes the value through, so the corrections
// alone give the same bits; otherwise the wave calls sqrt().
template <int NS>
__device__ __forceinline__ void sqrt_n(double* v) {
    bool ok = true;
#pragma unroll
    for (int k = 0; k < NS; k++) ok = ok && v[k] >= 0x1p-767 && v[k] < __builtin_inf();
    if (__all(ok)) {
#pragma unroll
        for (int k = 0; k < NS; k++) {
            const double x = v[k];
            const double y = __builtin_amdgcn_rsq(x);
            double g = x * y, h = 0.5 * y;
            const double r = fma(-h, g, 0.5);
            g = fma(g, r, g);
            h = fma(h, r, h);
            double d = fma(-g, g, x);
            g = fma(d, h, g);
            d = fma(-g, g, x);
            v[k] = fma(d, h, g);
        }
    } else {
#pragma unroll
        for (int k = 0; k < NS; k++) v[k] = sqrt(v[k]);
    }
}

// Volume term of inscribedSphereHex: |det| of the edges from node 0 to nodes 1, 3, 4 (geometry.cpp:308-355).
__device__ __forceinline__ double hex_corner_volume(const double (*x)[3]) {
#pragma clang fp contract(off)
    double M[9] = {x[1][0] - x[0][0], x[1][1] - x[0][1], x[1][2] - x[0][2],
                   x[3][0] - x[0][0], x[3][1] - x[0][1], x[3][2] - x[0][2],
                   x[4][0] - x[0][0], x[4][1] - x[0][1], x[4][2] - x[0][2]};
    return fabs(M[0] * (M[4] * M[8] - M[7] * M[5]) - M[1] * (M[3] * M[8] - M[6] * M[5]) +
                M[2] * (M[3] * M[7] - M[6] * M[4]));
}

// inscribedSphereHex (geometry.cpp:308-355) x2: classLinearHexahedra.cpp:209-234.
__device__ __forceinline__ double hex_char_length(const double (*x)[3]) {
#pragma clang fp contract(off)
    const int fc[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6}, {0, 3, 7, 4}};
    double n2[12];
#pragma unroll
    for (int f = 0; f < 6; f++) {
        n2[2 * f] = tri_norm2(x[fc[f][0]], x[fc[f][1]], x[fc[f][2]]);
        n2[2 * f + 1] = tri_norm2(x[fc[f][0]], x[fc[f][2]], x[fc[f][3]]);
    }
    sqrt_n<12>(n2);
    double A = 0.0;
#pragma unroll
    for (int f = 0; f < 6; f++) {
        double s = 0.5 * n2[2 * f] + 0.5 * n2[2 * f + 1];
        A = (f == 0) ? fabs(s) : A + fabs(s);
    }
    double V = hex_corner_volume(x);
    double r = 3.0 * V / A;
    return 2 * r;
}

// True when the element's stable dt min(l, l0)/c (l = hex_char_length(x)) is certainly above t, without
// the strict length.  Each face triangle (a,b,c), (a,c,d) of hex_char_length has area 1/2 |p x q| with p, q
// the two face edges at its quad corner (b, resp. d), and 1/2 |p x q| <= (|p|^2 + |q|^2)/4; every edge lies
// on two faces, so the area sum A <= S/2 with S the sum of the 12 squared edge lengths, and l = 6V/A >=
// 12V/S, V the same corner volume (bit for bit: hex_corner_volume).  The factor 1 + 2^-40 covers the
// rounding of A, S and the final division, all relative errors of a few ulp.
__device__ __forceinline__ bool hex_dt_above(const double (*x)[3], double l0, double c, double t) {
#pragma clang fp contract(off)
    const int ed[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6}, {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
    double S = 0.0;
#pragma unroll
    for (int k = 0; k < 12; k++) {
        const double d0 = x[ed[k][1]][0] - x[ed[k][0]][0], d1 = x[ed[k][1]][1] - x[ed[k][0]][1];
        const double d2 = x[ed[k][1]][2] - x[ed[k][0]][2];
        S += d0 * d0 + d1 * d1 + d2 * d2;
    }
    const double tc = t * c * (1.0 + 0x1p-40);
    return 12.0 * hex_corner_volume(x) > tc * S && l0 > tc;
}

// Volume term of inscribedSphere (geometry.cpp:264-301).
__device__ __forceinline__ double tet_volume(const double (*x)[3]) {
#pragma clang fp contract(off)
    double M[9] = {x[1][0] - x[0][0], x[1][1] - x[0][1], x[1][2] - x[0][2],
                   x[2][0] - x[0][0], x[2][1] - x[0][1], x[2][2] - x[0][2],
                   x[3][0] - x[0][0], x[3][1] - x[0][1], x[3][2] - x[0][2]};
    return 1.0 / 6.0 * fabs(M[0] * (M[4] * M[8] - M[7] * M[5]) - M[1] * (M[3] * M[8] - M[6] * M[5]) +
                            M[2] * (M[3] * M[7] - M[6] * M[4]));
}

// inscribedSphere (geometry.cpp:264-301) x2: classLinearTetrahedra.cpp:146-163.
__device__ __forceinline__ double tet_char_length(const double (*x)[3]) {
#pragma clang fp contract(off)
    double n2[4] = {tri_norm2(x[0], x[1], x[2]), tri_norm2(x[1], x[2], x[3]), tri_norm2(x[0], x[2], x[3]),
                    tri_norm2(x[0], x[1], x[3])};
    sqrt_n<4>(n2);
    double A = fabs(0.5 * n2[0]) + fabs(0.5 * n2[1]) + fabs(0.5 * n2[2]) + fabs(0.5 * n2[3]);
    double V = tet_volume(x);
    double r = 3.0 * V / A;
    return 2 * r;
}

// hex_dt_above for a tet4: each face's area is at most (|p|^2 + |q|^2)/4 for any two of its edges p, q, so at
// most a quarter of its two shorter squared edges; l = 6V/A >= 6V/A_ub with the same V (tet_volume).
__device__ __forceinline__ bool tet_dt_above(const double (*x)[3], double l0, double c, double t) {
#pragma clang fp contract(off)
    double q[6];  // squared edges 01 02 03 12 13 23
    const int ed[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const double d0 = x[ed[k][1]][0] - x[ed[k][0]][0], d1 = x[ed[k][1]][1] - x[ed[k][0]][1];
        const double d2 = x[ed[k][1]][2] - x[ed[k][0]][2];
        q[k] = d0 * d0 + d1 * d1 + d2 * d2;
    }
    // faces (0,1,2), (1,2,3), (0,2,3), (0,1,3): their edges' indices in q
    const int fe[4][3] = {{0, 1, 3}, {3, 4, 5}, {1, 2, 5}, {0, 2, 4}};
    double A4 = 0.0;  // 4 * the bound of the area sum
#pragma unroll
    for (int f = 0; f < 4; f++) {
        const double a = q[fe[f][0]], b = q[fe[f][1]], c2 = q[fe[f][2]];
        A4 += a + b + c2 - fmax(a, fmax(b, c2));
    }
    const double tc = t * c * (1.0 + 0x1p-40);
    return 24.0 * tet_volume(x) > tc * A4 && l0 > tc;
}

// FP32 characteristic lengths for the precision form (the reference formulas with FP32 areas, volume and
// ratio, hardware square root and reciprocal).  Tet: edge vectors differenced in FP64, so there is no
// cancellation against the absolute coordinates; hex: from node positions relative to their centroid.
__device__ __forceinline__ float area_tri_pf(const float* p, const float* q, const float* r) {
    const float A0 = p[0] - r[0], A1 = p[1] - r[1], A2 = p[2] - r[2];
    const float B0 = q[0] - r[0], B1 = q[1] - r[1], B2 = q[2] - r[2];
    const float C0 = A1 * B2 - A2 * B1, C1 = A2 * B0 - A0 * B2, C2 = A0 * B1 - A1 * B0;
    return 0.5f * __builtin_amdgcn_sqrtf(C0 * C0 + C1 * C1 + C2 * C2);  // hardware sqrt, 1 ulp
}

__device__ __forceinline__ float hex_char_length_pf(const float (*x)[3]) {
    const int fc[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6}, {0, 3, 7, 4}};
    float A = 0.0f;
#pragma unroll
    for (int f = 0; f < 6; f++)
        A += area_tri_pf(x[fc[f][0]], x[fc[f][1]], x[fc[f][2]]) + area_tri_pf(x[fc[f][0]], x[fc[f][2]], x[fc[f][3]]);
    float M[9];
    const int c[3] = {1, 3, 4};
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
        for (int d = 0; d < 3; d++) M[j * 3 + d] = x[c[j]][d] - x[0][d];
    return 6.0f * fabsf(det3(M)) * __builtin_amdgcn_rcpf(A);
}

// Tet: from the deformed edge vectors e_j = x_{j+1} - x_0 = Jm(:,j) + Du(:,j), whose parts the element already
// differenced in FP64: faces (0,1,2), (1,2,3), (0,2,3), (0,1,3) as |e1 x e2|, |(e2-e1) x (e3-e1)|, |e2 x e3|,
// |e1 x e3| over 2, and 6V = |det(e1, e2, e3)|.
__device__ __forceinline__ float cross_norm_f32(const float* a, const float* b) {
    const float c0 = a[1] * b[2] - a[2] * b[1], c1 = a[2] * b[0] - a[0] * b[2], c2 = a[0] * b[1] - a[1] * b[0];
    return __builtin_amdgcn_sqrtf(c0 * c0 + c1 * c1 + c2 * c2);  // hardware sqrt, 1 ulp
}

__device__ __forceinline__ float tet_char_length_f32(const float* Jm, const float* Du) {
    float e[3][3], f[2][3];
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
        for (int d = 0; d < 3; d++) e[j][d] = Jm[d * 3 + j] + Du[d * 3 + j];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        f[0][d] = e[1][d] - e[0][d];
        f[1][d] = e[2][d] - e[0][d];
    }
    const float A = 0.5f * (cross_norm_f32(e[0], e[1]) + cross_norm_f32(f[0], f[1]) + cross_norm_f32(e[1], e[2]) +
                            cross_norm_f32(e[0], e[2]));
    const float M[9] = {e[0][0], e[0][1], e[0][2], e[1][0], e[1][1], e[1][2], e[2][0], e[2][1], e[2][2]};
    return fabsf(det3(M)) * __builtin_amdgcn_rcpf(A);
}

// Element contribution to the stable dt: std::min(l0, l)/c (solvers.cpp:317).
__device__ __forceinline__ double elem_dt(double l, double l0, double c) {
#pragma clang fp contract(off)
    double m = (l < l0) ? l : l0;
    return m / c;
}

// ---------------------------------------------------------------------------
// Hex8 butterfly transform: 8 nodal vectors -> 7 derivative coefficient vectors
// A[0]=xi, A[1]=eta, A[2]=zeta, A[3]=xi*eta, A[4]=eta*zeta, A[5]=xi*zeta, A[6]=xi*eta*zeta,
// A[m] = sum_a sigma_m(a) x_a with sigma the product of the node signs.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void hex_butterfly(const double (*x)[3], double (*A)[3]) {
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double s0 = x[0][d] + x[1][d], d0 = x[1][d] - x[0][d];
        double s1 = x[3][d] + x[2][d], d1 = x[2][d] - x[3][d];
        double s2 = x[4][d] + x[5][d], d2 = x[5][d] - x[4][d];
        double s3 = x[7][d] + x[6][d], d3 = x[6][d] - x[7][d];
        double ssm = s0 + s1, sdm = s1 - s0, dsm = d0 + d1, ddm = d1 - d0;
        double ssp = s2 + s3, sdp = s3 - s2, dsp = d2 + d3, ddp = d3 - d2;
        A[2][d] = ssp - ssm;
        A[1][d] = sdm + sdp;
        A[4][d] = sdp - sdm;
        A[0][d] = dsm + dsp;
        A[5][d] = dsp - dsm;
        A[3][d] = ddm + ddp;
        A[6][d] = ddp - ddm;
    }
}

__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// The last two levels of the inverse transform from the pair sums and differences t.. (shared by both forms below).
template <typename T>
__device__ __forceinline__ void hex_unbutterfly_tail(T tssm, T tsdm, T tdsm, T tddm, T tssp, T tsdp, T tdsp, T tddp,
                                                     T (*f)[3], int d) {
#pragma clang fp contract(off)
    // zeta-, eta- (nodes 0,1)
    T us = tssm - tsdm, ud = tdsm - tddm;
    f[0][d] = us - ud;
    f[1][d] = us + ud;
    // zeta-, eta+ (nodes 3,2)
    us = tssm + tsdm;
    ud = tdsm + tddm;
    f[3][d] = us - ud;
    f[2][d] = us + ud;
    // zeta+, eta- (nodes 4,5)
    us = tssp - tsdp;
    ud = tdsp - tddp;
    f[4][d] = us - ud;
    f[5][d] = us + ud;
    // zeta+, eta+ (nodes 7,6)
    us = tssp + tsdp;
    ud = tdsp + tddp;
    f[7][d] = us - ud;
    f[6][d] = us + ud;
}

// Inverse transform with the coefficient scales folded in: f_a = sum_m sigma_m(a) s_m G[m] with s_m = s1 for the
// linear, s2 for the bilinear and s3 for the trilinear coefficient.  The products and sums are written out
// (explicit FMAs, no contraction), so the rounding does not depend on where the compiler places this code
// relative to other work of the kernel: the optional kinetic-energy block that follows it in the element kernel
// moved the scaling into another basic block, which changed which multiplies the compiler fused.
template <typename T>
__device__ __forceinline__ void hex_unbutterfly_scaled(const T (*G)[3], T s1, T s2, T s3, T (*f)[3]) {
#pragma clang fp contract(off)
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const T g2 = G[2][d] * s1, g4 = G[4][d] * s2, g5 = G[5][d] * s2, g6 = G[6][d] * s3;
        hex_unbutterfly_tail<T>(-g2, fma_t(G[1][d], s1, -g4), fma_t(G[0][d], s1, -g5), fma_t(G[3][d], s2, -g6), g2,
                                fma_t(G[1][d], s1, g4), fma_t(G[0][d], s1, g5), fma_t(G[3][d], s2, g6), f, d);
    }
}

// Inverse transform: f_a = sum_m sigma_m(a) G[m]  (G[m] already scaled), without contraction.
template <typename T>
__device__ __forceinline__ void hex_unbutterfly(const T (*G)[3], T (*f)[3]) {
#pragma clang fp contract(off)
#pragma unroll
    for (int d = 0; d < 3; d++)
        hex_unbutterfly_tail<T>(-G[2][d], G[1][d] - G[4][d], G[0][d] - G[5][d], G[3][d] - G[6][d], G[2][d],
                                G[1][d] + G[4][d], G[0][d] + G[5][d], G[3][d] + G[6][d], f, d);
}

// Scale the coefficient vectors into Jacobian units: linear terms by 1/8, bilinear by
// g/8, trilinear by g^2/8 (g = Gauss abscissa), converting to T.
template <typename T>
__device__ __forceinline__ void hex_scale(const double (*A)[3], T (*As)[3]) {
    const double g = MPS_GP;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        As[0][d] = (T)(0.125 * A[0][d]);
        As[1][d] = (T)(0.125 * A[1][d]);
        As[2][d] = (T)(0.125 * A[2][d]);
        As[3][d] = (T)((0.125 * g) * A[3][d]);
        As[4][d] = (T)((0.125 * g) * A[4][d]);
        As[5][d] = (T)((0.125 * g) * A[5][d]);
        As[6][d] = (T)((0.125 * g * g) * A[6][d]);
    }
}

// Sign factors of a Gauss point of the 2x2x2 rule, in the reference order (GPsDistribution.cpp:392-432:
// xi positive at points 1, 2, 5, 6, eta at 2, 3, 6, 7, zeta at 4-7), and their pairwise products, as +-1
// built from the point index by integer operations.  In a loop over the points the index is uniform, so
// the factors stay scalar operands of the FMAs instead of vector products.
template <typename T>
struct GpSigns {
    T x, y, z, yz, xz, xy;
};
__device__ __forceinline__ double unit_signed(double, unsigned neg) {
    return bits2d(0x3FF0000000000000ull | ((unsigned long long)neg << 63));
}
__device__ __forceinline__ float unit_signed(float, unsigned neg) { return __uint_as_float(0x3F800000u | (neg << 31)); }
template <typename T>
__device__ __forceinline__ GpSigns<T> gp_signs(int gp) {
    const unsigned nx = ((gp ^ (gp >> 1)) & 1) ^ 1, ny = ((gp >> 1) & 1) ^ 1, nz = ((gp >> 2) & 1) ^ 1;
    return {unit_signed(T(), nx),      unit_signed(T(), ny),      unit_signed(T(), nz),
            unit_signed(T(), ny ^ nz), unit_signed(T(), nx ^ nz), unit_signed(T(), nx ^ ny)};
}

// Material parameters of element e: (pa, pb) = (lambda, mu) [compressible] or (mu, K1) [Belytschko].
__device__ __forceinline__ void elem_mat(const MpsInst& I, int e, double& pa, double& pb, double& c) {
    if (I.mat) {
        pa = I.mat[2 * (size_t)e];
        pb = I.mat[2 * (size_t)e + 1];
        c = I.cel[e];
    } else {
        pa = I.mat_a;
        pb = I.mat_b;
        c = I.c0;
    }
}

// ---------------------------------------------------------------------------
// Bulk artificial viscosity at one Gauss point (bulkViscosity, solverExplicit.cpp:513-542):
// trD = tr(F1 - F0)/dt, q = rho/J h (c1^2 h |trD| min(0, trD) + cl c0 trD) and the PK1
// -J (-q I) F^-T = q cof(F), added to P.  qa = rho h c1^2 h and qb = rho h cl c0 per element.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void add_bulk_visc(const T* H, T trD, T qa, T qb, T* P) {
    const T J = T(1) + detIpH_m1(H);
    const T mn = trD < T(0) ? trD : T(0);
    const T q = (qa * fabs(trD) * mn + qb * trD) / J;
    T F[9], cF[9];
#pragma unroll
    for (int i = 0; i < 9; i++) F[i] = H[i];
    F[0] += T(1);
    F[4] += T(1);
    F[8] += T(1);
    cof3(F, cF);
#pragma unroll
    for (int i = 0; i < 9; i++) P[i] += q * cF[i];
}

// Viscosity coefficients of element e (qa, qb of add_bulk_visc); element size and sound speed are the
// element's undeformed length and sound speed unless given.
__device__ __forceinline__ void visc_coeffs(const MpsInst& I, int e, double c, double* qa, double* qb) {
    const double rho = I.rho ? I.rho[e] : I.rho0;
    const double h = I.visc_h > 0 ? I.visc_h : I.l0[e];
    const double c0 = I.visc_c0 > 0 ? I.visc_c0 : c;
    *qa = rho * h * I.visc_c1 * I.visc_c1 * h;
    *qb = rho * h * I.visc_cl * c0;
}

// Gather of nodal 3-vectors stored by direction (p[d*N + n]) at the nodes nd[0..NN-1], from one array (P) or,
// with TWO, from two (P and Q).  O32 (meshes of at most MPS_OFF32_NODES nodes): a node's value is addressed by a
// 32-bit byte offset from the direction's base address, a scalar, so that one shift per node serves every
// direction and array; otherwise by 64-bit addresses (one 64-bit add per value).
template <int NN, bool O32, bool TWO = false>
__device__ __forceinline__ void gather_nodes(size_t N, const int* nd, const double* p, double (*P)[3],
                                             const double* q = nullptr, double (*Q)[3] = nullptr) {
    if constexpr (O32) {
        uint32_t off[NN];
#pragma unroll
        for (int a = 0; a < NN; a++) off[a] = (uint32_t)nd[a] * 8u;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const char* bp = reinterpret_cast<const char*>(p + d * N);
            const char* bq = reinterpret_cast<const char*>(q + d * N);
#pragma unroll
            for (int a = 0; a < NN; a++) {
                P[a][d] = *reinterpret_cast<const double*>(bp + off[a]);
                if constexpr (TWO) Q[a][d] = *reinterpret_cast<const double*>(bq + off[a]);
            }
        }
    } else {
#pragma unroll
        for (int a = 0; a < NN; a++)
#pragma unroll
            for (int d = 0; d < 3; d++) {
                P[a][d] = p[d * N + nd[a]];
                if constexpr (TWO) Q[a][d] = q[d * N + nd[a]];
            }
    }
}

// Consistent kinetic energy of hex8 element e (KIN_ENERGY, dataExtraction.cpp:631-707): sum over the GPs of
// wJ[gp] rho/2 |v(gp)|^2, v(gp) = sum_a phi_a(gp) v_a evaluated from the butterfly coefficients of the nodal
// velocities, (sum_a v_a + sum_m s_m(gp) g^|m| C_m) / 8.
template <bool O32 = false>
__device__ __forceinline__ double hex_kinetic(const MpsInst& I, const int* nd, double rho, const double* wJ) {
    const size_t N = (size_t)I.N;
    double x[8][3], Cv[7][3], vs[3];
    gather_nodes<8, O32>(N, nd, I.v, x);
    hex_butterfly(x, Cv);
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double t = 0.0;
#pragma unroll
        for (int a = 0; a < 8; a++) t += x[a][d];
        vs[d] = t;
    }
    const double g = MPS_GP, g2 = MPS_GP * MPS_GP, g3 = MPS_GP * MPS_GP * MPS_GP;
    double ek = 0.0;
#pragma unroll
    for (int gp = 0; gp < 8; gp++) {
        const int sx = sxi(gp), sy = seta(gp), sz = szeta(gp);
        double kv = 0.0;
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const double vel = 0.125 * (vs[j] + (sx * g) * Cv[0][j] + (sy * g) * Cv[1][j] + (sz * g) * Cv[2][j] +
                                        (sx * sy * g2) * Cv[3][j] + (sy * sz * g2) * Cv[4][j] +
                                        (sx * sz * g2) * Cv[5][j] + (sx * sy * sz * g3) * Cv[6][j]);
            kv += 0.5 * rho * vel * vel;
        }
        ek += kv * wJ[gp];
    }
    return ek;
}

// Jacobian (or displacement gradient) D(i,j) = d(.)_i/d xi_j at the Gauss point with signs (sx, sy, sz) from the
// scaled butterfly coefficients A (hex_scale).
template <typename T>
__device__ __forceinline__ void hex_jac(const T (*A)[3], const GpSigns<T>& s, T* D) {
#pragma unroll
    for (int i = 0; i < 3; i++) {
        D[i * 3 + 0] = A[0][i] + s.y * A[3][i] + s.z * A[5][i] + s.yz * A[6][i];
        D[i * 3 + 1] = A[1][i] + s.x * A[3][i] + s.z * A[4][i] + s.xz * A[6][i];
        D[i * 3 + 2] = A[2][i] + s.x * A[5][i] + s.y * A[4][i] + s.xy * A[6][i];
    }
}

// ---------------------------------------------------------------------------
// Hex8 element: nodal forces f[a][i] (T) and the element dt contribution (FP64).
// VISC adds bulk viscosity from the displacement increment u - uprev over the step (rdt = 1/dt, 0 when
// dt = 0); its 7 butterfly coefficient vectors wait in the element's LDS force slots vs[m*3+d] (stride
// MPS_PATCH), which the forces overwrite only after the Gauss-point loop.
// EN (FP64, inviscid): DEFO_ENERGY and VOLUME of u in en[0], en[1] (sum over the GPs of wJ W(F), wJ det F); the
// GPs' wJ go to vs[gp*MPS_PATCH] (an LDS array apart from the force slots), for the kinetic energy, which the
// element kernel computes after the force phase (hex_kinetic).
// ---------------------------------------------------------------------------
// dtsk (FP64): the element kernel's dt skip threshold (MpsCtl::dt_skip; +inf: none).  A wave whose elements all
// have hex_dt_above(dtsk) returns +inf as their dt and skips the strict length.
template <typename T, int LAW, bool VISC = false, bool EN = false, bool O32 = false>
__device__ __forceinline__ double hex8_element(const MpsInst& I, int e, const double* __restrict__ u, T (*f)[3],
                                               bool want_dt, T* vs = nullptr, double rdt = 0.0,
                                               unsigned eflags = 0u, double* en = nullptr,
                                               double dtsk = __builtin_inf()) {
    static_assert(!EN || (sizeof(T) == 8 && !VISC), "fused energies: FP64 inviscid element kernels");
    double e_def = 0.0, e_vol = 0.0;
    const int32_t* cn = I.conn + (size_t)e * 8;
    int4 q0 = *reinterpret_cast<const int4*>(cn);
    int4 q1 = *reinterpret_cast<const int4*>(cn + 4);
    const int nd[8] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w};
    const size_t N = (size_t)I.N;
    double pa, pb, c;
    elem_mat(I, e, pa, pb, c);
    T Ax[7][3], Au[7][3], G[7][3];
    double dte = bits2d(0x7FF0000000000000ull);
    {
        double xX[8][3], xU[8][3], cx[7][3];
        gather_nodes<8, O32, true>(N, nd, I.X, xX, u, xU);
        if (want_dt && sizeof(T) == 8) {
            // deformed characteristic length (solvers.cpp:300-317), strict FP64: bitwise the reference's.
            // Taken here, while the gathered nodes are in registers: after the Gauss-point loop the
            // nodes would have to be read again, with their latency exposed.
            double x[8][3];
#pragma unroll
            for (int a = 0; a < 8; a++)
#pragma unroll
                for (int d = 0; d < 3; d++) x[a][d] = xX[a][d] + xU[a][d];
            const double l0 = I.l0[e];
            // dtsk is the same in every lane: a uniform branch
            if (!(dtsk < __builtin_inf()) || __any(!hex_dt_above(x, l0, c, dtsk))) dte = elem_dt(hex_char_length(x), l0, c);
        }
        hex_butterfly(xX, cx);
        hex_scale<T>(cx, Ax);
        hex_butterfly(xU, cx);
        hex_scale<T>(cx, Au);
    }
    T qa = T(0), qb = T(0), trs = T(0);
    if (VISC) {
        double x[8][3], cx[7][3];
        T dA[7][3];
#pragma unroll
        for (int a = 0; a < 8; a++)
#pragma unroll
            for (int d = 0; d < 3; d++) x[a][d] = u[d * N + nd[a]] - I.uprev[d * N + nd[a]];
        hex_butterfly(x, cx);
        hex_scale<T>(cx, dA);
#pragma unroll
        for (int m = 0; m < 7; m++)
#pragma unroll
            for (int d = 0; d < 3; d++) vs[(m * 3 + d) * MPS_PATCH] = dA[m][d];
        double da, db;
        visc_coeffs(I, e, c, &da, &db);
        qa = (T)da;
        qb = (T)db;
        trs = (T)rdt;
    }
#pragma unroll
    for (int m = 0; m < 7; m++)
#pragma unroll
        for (int d = 0; d < 3; d++) G[m][d] = T(0);
    const T tpa = (T)pa, tpb = (T)pb;
#pragma unroll 1
    for (int gp = 0; gp < 8; gp++) {
        // Gauss point signs in the reference order (GPsDistribution.cpp:392-432)
        const GpSigns<T> s = gp_signs<T>(gp);
        T Jm[9], Du[9], H[9], P[9];
        hex_jac<T>(Ax, s, Jm);
        hex_jac<T>(Au, s, Du);
        // J^-1 = cof(J)^T / det J ; weight 1 (GPsDistribution.cpp:392-432) so wJ J^-T = sign(det J) cof(J)
        T cJ[9];
        cof3(Jm, cJ);
        const T dJ = Jm[0] * cJ[0] + Jm[1] * cJ[1] + Jm[2] * cJ[2];
        const T rJ = rcp(dJ);
        // H = Du J^-1: H(i,j) = sum_k Du(i,k) cof(J)(j,k) / det J
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++)
                H[i * 3 + j] = (Du[i * 3] * cJ[j * 3] + Du[i * 3 + 1] * cJ[j * 3 + 1] + Du[i * 3 + 2] * cJ[j * 3 + 2]) * rJ;
        if constexpr (EN) {  // weight 1: wJ = |det J| (classGPs.cpp:26-27)
            double Wg, Jg;
            pk1<LAW>(H, tpa, tpb, P, &Wg, &Jg);
            const double wJ = fabs(dJ);
            e_def += Wg * wJ;
            e_vol += wJ * Jg;
            vs[gp * MPS_PATCH] = wJ;
        } else {
            pk1<LAW>(H, tpa, tpb, P);
        }
        if (VISC) {  // tr of grad(u - uprev) = sum_ik Dd(i,k) cof(J)(i,k) / det J
            // the memory clobber keeps the 21 LDS reads at each Gauss point: hoisted out of the loop
            // they would hold 21 more values in registers throughout it.  The FP64 Belytschko kernel
            // runs one wave per SIMD either way and is faster with them hoisted.
            if (!(sizeof(T) == 8 && LAW == 1)) asm volatile("" ::: "memory");
            T tr = T(0);
#pragma unroll
            for (int i = 0; i < 3; i++) {
                const T a0 = vs[(0 * 3 + i) * MPS_PATCH], a1 = vs[(1 * 3 + i) * MPS_PATCH], a2 = vs[(2 * 3 + i) * MPS_PATCH];
                const T a3 = vs[(3 * 3 + i) * MPS_PATCH], a4 = vs[(4 * 3 + i) * MPS_PATCH], a5 = vs[(5 * 3 + i) * MPS_PATCH];
                const T a6 = vs[(6 * 3 + i) * MPS_PATCH];
                tr += (a0 + s.y * a3 + s.z * a5 + s.yz * a6) * cJ[i * 3 + 0] +
                      (a1 + s.x * a3 + s.z * a4 + s.xz * a6) * cJ[i * 3 + 1] +
                      (a2 + s.x * a5 + s.y * a4 + s.xy * a6) * cJ[i * 3 + 2];
            }
            add_bulk_visc<T>(H, tr * rJ * trs, qa, qb, P);
        }
        // Q = wJ P J^-T: Q(i,k) = sign(det J) sum_J P(i,J) cof(J)(J,k); columns Q(:,k).  The sign is
        // folded into P only when some lane of the wave has det J <= 0 (a positively oriented
        // mesh never takes that branch).
        if (!__all(dJ > T(0))) {
            const T sg = dJ < T(0) ? T(-1) : T(1);
#pragma unroll
            for (int i = 0; i < 9; i++) P[i] *= sg;
        }
        T Q[3][3];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int k = 0; k < 3; k++) Q[k][i] = P[i * 3] * cJ[k] + P[i * 3 + 1] * cJ[3 + k] + P[i * 3 + 2] * cJ[6 + k];
        // moment sums, one FMA per (moment, column) pair: a +-1 factor makes each product exact
#pragma unroll
        for (int i = 0; i < 3; i++) {
            G[0][i] += Q[0][i];
            G[1][i] += Q[1][i];
            G[2][i] += Q[2][i];
            G[3][i] = fma_t(s.x, Q[1][i], fma_t(s.y, Q[0][i], G[3][i]));
            G[4][i] = fma_t(s.y, Q[2][i], fma_t(s.z, Q[1][i], G[4][i]));
            G[5][i] = fma_t(s.x, Q[2][i], fma_t(s.z, Q[0][i], G[5][i]));
            G[6][i] = fma_t(s.xy, Q[2][i], fma_t(s.xz, Q[1][i], fma_t(s.yz, Q[0][i], G[6][i])));
        }
    }
    hex_unbutterfly_scaled<T>(G, (T)0.125, (T)(0.125 * MPS_GP), (T)(0.125 * MPS_GP * MPS_GP), f);
    if constexpr (EN) {
        en[0] = e_def;
        en[1] = e_vol;
    }
    if (want_dt && sizeof(T) == 4) {
        // precision form: FP32 length from the deformed node positions relative to their centroid,
        // x_a - xbar = sum_m sigma_m(a) C_m / 8 with C the butterfly coefficients of X + u (Ax + Au
        // unscaled), so no node data is read again; dt and its reduction stay FP64
        const T ig = (T)(1.0 / MPS_GP), ig2 = (T)(1.0 / (MPS_GP * MPS_GP));
        T R[7][3], xr[8][3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            R[0][d] = Ax[0][d] + Au[0][d];
            R[1][d] = Ax[1][d] + Au[1][d];
            R[2][d] = Ax[2][d] + Au[2][d];
            R[3][d] = (Ax[3][d] + Au[3][d]) * ig;
            R[4][d] = (Ax[4][d] + Au[4][d]) * ig;
            R[5][d] = (Ax[5][d] + Au[5][d]) * ig;
            R[6][d] = (Ax[6][d] + Au[6][d]) * ig2;
        }
        hex_unbutterfly<T>(R, xr);
        dte = elem_dt((double)hex_char_length_pf((const float(*)[3])xr), I.l0[e], c);
    }
    return dte;
}

// ---------------------------------------------------------------------------
// Tet4 element (1 Gauss point): classLinearTetrahedra.cpp:171-292.
// J(i,j) = x_{j+1,i} - x_{0,i}; dphi_0 = -(sum of J^-1 rows), dphi_a = row a-1 of J^-1.
// ---------------------------------------------------------------------------
// Per-element constants of the tet kernels: law parameters, sound speed and undeformed length (read
// only when the element's dt is wanted).  The empty asm makes c a register value, so that the compiler
// does not turn the uniform/per-element choice into a pointer select with a private copy of the kernel
// argument.
struct TetConsts {
    double pa, pb, c, l0;
};
__device__ __forceinline__ TetConsts tet_consts(const MpsInst& I, int e, bool want_dt) {
    TetConsts k;
    elem_mat(I, e, k.pa, k.pb, k.c);
    asm volatile("" : "+v"(k.c));
    k.l0 = want_dt ? I.l0[e] : 0.0;
    return k;
}

// tet4_core takes the element's node coordinates x, displacements uu and constants; nd (node ids) is
// read only by the viscous form.  tet4_element gathers them from X, u and the material arrays.
// EN (FP64, inviscid): en[0], en[1] = wJ W(F), wJ det F of the element's Gauss point, en[3] = its wJ (the
// kinetic energy needs the node ids, which the callers hold).
template <typename T, int LAW, bool VISC = false, bool EN = false>
__device__ __forceinline__ double tet4_core(const MpsInst& I, int e, const int* nd, double (*x)[3], const double (*uu)[3],
                                            const TetConsts& k, T (*f)[3], bool want_dt, double rdt,
                                            double* en = nullptr) {
    static_assert(!EN || (sizeof(T) == 8 && !VISC), "fused energies: FP64 inviscid element kernels");
    const size_t N = (size_t)I.N;
    const double pa = k.pa, pb = k.pb, c = k.c;
    T Jm[9], Du[9];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            Jm[i * 3 + j] = (T)(x[j + 1][i] - x[0][i]);
            Du[i * 3 + j] = (T)(uu[j + 1][i] - uu[0][i]);
        }
    double dte = bits2d(0x7FF0000000000000ull);
    if (want_dt) {
        double l;
        if constexpr (sizeof(T) == 8) {
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int d = 0; d < 3; d++) x[a][d] = x[a][d] + uu[a][d];
            l = tet_char_length(x);
        } else {
            l = (double)tet_char_length_f32(Jm, Du);
        }
        dte = elem_dt(l, k.l0, c);
    }
    T iJ[9], H[9], P[9];
    T dJ = det3(Jm);
    inv3(Jm, iJ, dJ);
    T wJ = (T)MPS_TET_W * fabs(dJ);
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) H[i * 3 + j] = Du[i * 3] * iJ[j] + Du[i * 3 + 1] * iJ[3 + j] + Du[i * 3 + 2] * iJ[6 + j];
    if constexpr (EN) {
        double Wg, Jg;
        pk1<LAW>(H, pa, pb, P, &Wg, &Jg);
        en[0] = Wg * wJ;
        en[1] = wJ * Jg;
        en[3] = wJ;
    } else {
        pk1<LAW>(H, (T)pa, (T)pb, P);
    }
    if (VISC) {  // tr of grad(u - uprev) = sum_ik Dd(i,k) J^-1(k,i)
        T Dd[9];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++)
                Dd[i * 3 + j] = (T)((uu[j + 1][i] - I.uprev[i * N + nd[j + 1]]) - (uu[0][i] - I.uprev[i * N + nd[0]]));
        T tr = T(0);
#pragma unroll
        for (int i = 0; i < 3; i++) tr += Dd[i * 3] * iJ[i] + Dd[i * 3 + 1] * iJ[3 + i] + Dd[i * 3 + 2] * iJ[6 + i];
        double da, db;
        visc_coeffs(I, e, c, &da, &db);
        add_bulk_visc<T>(H, tr * (T)rdt, (T)da, (T)db, P);
    }
    T dphi[4][3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        dphi[0][j] = -iJ[j] - iJ[3 + j] - iJ[6 + j];
        dphi[1][j] = iJ[j];
        dphi[2][j] = iJ[3 + j];
        dphi[3][j] = iJ[6 + j];
    }
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int i = 0; i < 3; i++)
            f[a][i] = wJ * P[i * 3] * dphi[a][0] + wJ * P[i * 3 + 1] * dphi[a][1] + wJ * P[i * 3 + 2] * dphi[a][2];
    return dte;
}

// Kinetic energy of tet4 element e at its Gauss point (dataExtraction.cpp:631-707): wJ rho/2 |v(gp)|^2 with
// v(gp) = sum_a v_a / 4 (the shape values at (1/4, 1/4, 1/4)).
__device__ __forceinline__ double tet_kinetic(const MpsInst& I, const int* nd, double rho, double wJ) {
    const size_t N = (size_t)I.N;
    double vel[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int j = 0; j < 3; j++) vel[j] += 0.25 * I.v[j * N + nd[a]];
    double kv = 0.0;
#pragma unroll
    for (int j = 0; j < 3; j++) kv += 0.5 * rho * vel[j] * vel[j];
    return kv * wJ;
}

template <typename T, int LAW, bool VISC = false, bool EN = false, bool O32 = false>
__device__ __forceinline__ double tet4_element(const MpsInst& I, int e, const double* __restrict__ u, T (*f)[3],
                                               bool want_dt, double rdt = 0.0, unsigned eflags = 0u,
                                               double* en = nullptr) {
    const int4 q = *reinterpret_cast<const int4*>(I.conn + (size_t)e * 4);
    const int nd[4] = {q.x, q.y, q.z, q.w};
    const size_t N = (size_t)I.N;
    double x[4][3], uu[4][3];
    gather_nodes<4, O32, true>(N, nd, I.X, x, u, uu);
    if constexpr (EN) {
        double r[4];
        const double dte = tet4_core<T, LAW, false, true>(I, e, nd, x, uu, tet_consts(I, e, want_dt), f, want_dt, rdt, r);
        en[0] = r[0];
        en[1] = r[1];
        en[2] = (eflags & MPS_E_KIN) ? tet_kinetic(I, nd, I.rho ? I.rho[e] : I.rho0, r[3]) : 0.0;
        return dte;
    } else {
        return tet4_core<T, LAW, VISC>(I, e, nd, x, uu, tet_consts(I, e, want_dt), f, want_dt, rdt);
    }
}

// ---------------------------------------------------------------------------
// Step control (begin of step k): lagged time update, termination, stable dt.
// solverExplicit.cpp:217-273.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int begin_scalars(const MpsInst& I, long long k, const MpsStepSlot& prev,
                                             unsigned long long dtmin_bits, MpsStepSlot* out) {
#pragma clang fp contract(off)
    out->go = 0;
    out->time = prev.time;
    out->dt = prev.dt;
    if (!prev.go) return MPSD_RUNNING;
    double time;
    if (I.time_mode == MPS_TIME_FIXED) {
        if (k > I.n_steps) return MPSD_DONE;
        time = prev.time + prev.dt;
    } else {
        if (!(prev.time < I.tf - 1E-12)) return MPSD_DONE;
        if (prev.time + prev.dt > I.tf) time = I.tf;
        else time = prev.time + prev.dt;
    }
    double dt = bits2d(dtmin_bits);
    dt = dt * I.scale;
    if (dt < 1E-20) return MPSD_ERR_DT;
    out->time = time;
    out->dt = dt;
    out->go = 1;
    return MPSD_RUNNING;
}

// Stable-dt skip threshold for the element kernel of step k, from the min dt of step k-1 (bit pattern mb, before
// the scale factor): elements certainly above it skip the strict length (hex_dt_above), and the kernel's last
// workgroup checks that the step's min is at most it (dt_certify).  +inf: every element exact.
__device__ __forceinline__ double dt_skip_next(const MpsInst& I, long long k, unsigned long long mb) {
    if (!I.dt_skip_on || k < I.ctl->skip_hold) return __builtin_inf();
    return bits2d(mb) * (1.0 + I.dt_skip_margin);
}

// Prescribed motion of a constrained DOF at step k with step dt (applyDisplacementBC, solverExplicit.cpp:545-575;
// classNodes.cpp:58-74): increment U and velocity V.  Ramp (val = UU = U/duration): U = V = UU dt; instantaneous:
// U = V = val at the first step, 0 after; velocity amplitude (extension): V = val min(k/ramp, 1), U = V dt.
__device__ __forceinline__ void bc_motion(const MpsInst& I, double val, long long k, double dt, double* U, double* V) {
#pragma clang fp contract(off)
    if (I.bc_mode == 1) {
        *U = (k == 1) ? val : 0.0;
        *V = *U;
    } else if (I.bc_mode == 0) {
        *U = val * dt;
        *V = *U;
    } else {
        double fr = (double)k / (double)(I.vel_ramp > 0 ? I.vel_ramp : 1);
        if (fr > 1.0) fr = 1.0;
        *V = val * fr;
        *U = *V * dt;
    }
}

// Half-step velocity w = v + dt/2 a of a DOF.  The central-difference step of the reference
// (nonlinearSystem.cpp:1186-1201, 1096-1107) is, with the dt of the step,
//   u_{k+1} = u_k + dt v_k + dt^2/2 a_k,   v_{k+1} = v_k + dt (a_k + a_{k+1}) / 2,
// which is u_{k+1} = u_k + dt w_k and v_{k+1} = w_k + dt/2 a_{k+1} with w_k = v_k + dt/2 a_k: between two steps
// of one launch only w has to be kept (the node kernel reads and writes 24 bytes per node less than with v
// and a).  Constrained DOFs hold a = 0, so there w = v.
__device__ __forceinline__ double half_velocity(double dt, double v, double a) { return __builtin_fma(0.5 * dt, a, v); }

// Increment of DOF i of node n at the begin of step k: the prediction dt w on active DOFs (w the half-step
// velocity above), the prescribed U on constrained ones.
__device__ __forceinline__ double step_increment(const MpsInst& I, uint32_t info, int i, long long k, double dt,
                                                 double w) {
#pragma clang fp contract(off)
    if (!((info >> i) & 1u)) return dt * w;
    double U, V;
    bc_motion(I, I.bcval[3 * (size_t)((info & MPS_REC_MASK) >> 3) + i], k, dt, &U, &V);
    return U;
}

// Prediction + Dirichlet displacement of node n for step k (nonlinearSystem.cpp:1186-1201,
// solverExplicit.cpp:545-575), from the half-step velocity w.  The prescribed velocity and the zero
// acceleration of the constrained DOFs are set by the solve of the same step (node_update).
__device__ __forceinline__ void predict_node(const MpsInst& I, int n, long long k, double dt, double* u,
                                             const double* w) {
    const uint32_t info = I.ninfo[n];
#pragma unroll
    for (int i = 0; i < 3; i++) u[i] += step_increment(I, info, i, k, dt, w[i]);
}

// External force of flagged node j (lnn[j]) at time t (NeumannBCManagement, boundaryConditions.cpp:75-140):
// zero, then its load items in input order -- nodal force entries (BCsNodalForces :1120-1134 with
// classNeumann::getTracAmount, classNodes.h:101-108: F/(tf-t0) * t for t0 <= t <= tf, else 0), the step's
// surface forces of pressure faces, and its gravity load (BCsGravity :1146-1172, constant, summed over the
// Gauss points at set-up) at the place of *GRAVITY.
__device__ __forceinline__ void loaded_node_fext(const MpsInst& I, int j, double t, double* fe) {
#pragma clang fp contract(off)
    fe[0] = 0.0;
    fe[1] = 0.0;
    fe[2] = 0.0;
    const int n = I.lnn[j];
    for (int q = I.lrow[j]; q < I.lrow[j + 1]; q++) {
        const uint32_t code = (uint32_t)I.lcode[q];
        const uint32_t kind = code >> 30;
        const size_t idx = code & MPS_ITEM_INDEX;
        if (kind == MPS_ITEM_FORCE) {
            const bool in = t >= I.lwin[2 * idx] && t <= I.lwin[2 * idx + 1];
            for (int d = 0; d < 3; d++) fe[d] += in ? I.lm[3 * idx + d] * t : 0.0;
        } else if (kind == MPS_ITEM_PRESSURE) {
            for (int d = 0; d < 3; d++) fe[d] += I.pc[3 * idx + d];
        } else {
            fe[2] += I.fgrav ? I.fgrav[n] : 0.0;
        }
    }
}

// External force of node n in the node kernel: the flagged node's value of this step (lf), else its
// gravity load alone (zero without gravity).
__device__ __forceinline__ void node_fext(const MpsInst& I, int n, uint32_t info, double* fe) {
    if (info & MPS_NODE_LOADED) {
        const double* f = I.lf + 3 * (size_t)I.lidx[n];
        fe[0] = f[0];
        fe[1] = f[1];
        fe[2] = f[2];
    } else {
        fe[0] = 0.0;
        fe[1] = 0.0;
        fe[2] = I.fgrav ? I.fgrav[n] : 0.0;
    }
}

// Surface forces of pressure Gauss point q at time t (BCsPressureRamp / BCsPressureInst,
// boundaryConditions.cpp:370-496): F of the bulk element at the point (classSpatialPoint::getF,
// classSpatialPoint.cpp:41-68: node terms first, then the identity), J = det F, n = F^-T n0, and
// pc[(q*8 + a)*3 + i] = w|J_s| phi_a J P(t) n_i for the scatter nodes a.  FP64, reference operation order.
__device__ __forceinline__ void surface_point(const MpsInst& I, int q, double t) {
#pragma clang fp contract(off)
    const size_t N = (size_t)I.N;
    const int nen = I.nen;
    const int e = I.sge[q];
    const double* D = I.sgd + (size_t)q * MPS_SGP_STRIDE;
    const double* B = I.sblk + MPS_SBLK_STRIDE * (size_t)I.sgb[q];
    double* out = I.pc + (size_t)q * 24;
    if (B[0] >= 4.0) {  // traction (BCsTractionRamp / Inst, boundaryConditions.cpp:218-271): w|J_s| phi_a T_i
        const bool tramp = B[0] == 4.0;
        double T[3] = {0., 0., 0.};
        if (t >= B[2] && t <= B[3])
            for (int i = 0; i < 3; i++) T[i] = tramp ? B[4 + i] / (B[3] - B[2]) * (t - B[2]) : B[4 + i];
        for (int a = 0; a < 8; a++)
            for (int i = 0; i < 3; i++) out[a * 3 + i] = D[27] * D[28 + a] * T[i];
        return;
    }
    double F[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int a = 0; a < nen; a++) {
        const int nd = I.conn[(size_t)e * nen + a];
        double disp[3];
        for (int i = 0; i < 3; i++) disp[i] = I.u[i * N + nd];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) F[i * 3 + j] += D[a * 3 + j] * disp[i];
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) F[i * 3 + j] += (i == j) ? 1.0 : 0.0;
    const double J = det3_strict(F);
    const double idet = 1.0 / J;
    double Fi[9];
    Fi[0] = idet * (F[4] * F[8] - F[5] * F[7]);
    Fi[3] = -idet * (F[3] * F[8] - F[5] * F[6]);
    Fi[6] = idet * (F[3] * F[7] - F[4] * F[6]);
    Fi[1] = -idet * (F[1] * F[8] - F[2] * F[7]);
    Fi[4] = idet * (F[0] * F[8] - F[2] * F[6]);
    Fi[7] = -idet * (F[0] * F[7] - F[1] * F[6]);
    Fi[2] = idet * (F[1] * F[5] - F[2] * F[4]);
    Fi[5] = -idet * (F[0] * F[5] - F[2] * F[3]);
    Fi[8] = idet * (F[0] * F[4] - F[1] * F[3]);
    const double* n0 = D + 24;
    double nv[3];
    for (int i = 0; i < 3; i++) {
        double r = 0.0;
        for (int j = 0; j < 3; j++) r += Fi[j * 3 + i] * n0[j];  // (F^-1)^T, row i
        nv[i] = r;
    }
    const bool ramp = B[0] == 2.0;
    double P = 0.;
    if (t >= B[2] && t <= B[3]) P = ramp ? B[1] / (B[3] - B[2]) * (t - B[2]) : B[1];
    const double ws = D[27];
    for (int a = 0; a < 8; a++) {
        const double ph = D[28 + a];
        for (int i = 0; i < 3; i++) out[a * 3 + i] = ws * ph * J * P * nv[i];
    }
}

// Partial nodal force g (FP64 partials; FP32 in the precision form).
__device__ __forceinline__ void load_partial(const MpsInst& I, int g, double* c) {
    if (I.precision == 1) {
        const float* p = reinterpret_cast<const float*>(I.part) + 3 * (size_t)g;
        c[0] = p[0];
        c[1] = p[1];
        c[2] = p[2];
    } else {
        const double* p = I.part + 3 * (size_t)g;
        c[0] = p[0];
        c[1] = p[1];
        c[2] = p[2];
    }
}

// Internal force of node n: the sum of its patch partials in ascending patch order, in FP64 (the precision
// form's FP32 partials widened first).
__device__ __forceinline__ void node_partial_sum(const MpsInst& I, int n, double* f) {
    int q = I.nrow[n];
    const int q1 = I.nrow[n + 1];
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    if (I.precision == 1) {  // FP32 partials of the precision form, summed in FP64
        const float* __restrict__ part = reinterpret_cast<const float*>(I.part);
        for (; q + 2 <= q1; q += 2) {
            const float* c0 = part + 3 * (size_t)I.ncol[q];
            const float* c1 = part + 3 * (size_t)I.ncol[q + 1];
            const double a0 = c0[0], b0 = c0[1], d0 = c0[2];
            const double a1 = c1[0], b1 = c1[1], d1 = c1[2];
            f0 += a0; f1 += b0; f2 += d0;
            f0 += a1; f1 += b1; f2 += d1;
        }
        if (q < q1) {
            const float* c = part + 3 * (size_t)I.ncol[q];
            f0 += (double)c[0];
            f1 += (double)c[1];
            f2 += (double)c[2];
        }
    } else {
        const double* __restrict__ part = I.part;
        // 2 entries per iteration: both loads issue before the (ordered) additions
        for (; q + 2 <= q1; q += 2) {
            const double* c0 = part + 3 * (size_t)I.ncol[q];
            const double* c1 = part + 3 * (size_t)I.ncol[q + 1];
            const double a0 = c0[0], b0 = c0[1], d0 = c0[2];
            const double a1 = c1[0], b1 = c1[1], d1 = c1[2];
            f0 += a0; f1 += b0; f2 += d0;
            f0 += a1; f1 += b1; f2 += d1;
        }
        if (q < q1) {
            const double* c = part + 3 * (size_t)I.ncol[q];
            f0 += c[0];
            f1 += c[1];
            f2 += c[2];
        }
    }
    f[0] = f0;
    f[1] = f1;
    f[2] = f2;
}

// rhs = -fint + fext (nonlinearSystem.cpp:1096-1107).
__device__ __forceinline__ void node_rhs(const double* fi, const double* fe, double* rhs) {
#pragma unroll
    for (int i = 0; i < 3; i++) rhs[i] = -fi[i] + fe[i];
}

// Velocity and acceleration of node n at the end of step k: the prescribed velocity and zero acceleration of
// its constrained DOFs (applyDisplacementBC, solverExplicit.cpp:568-571), the lumped-mass solve of the free
// ones (nonlinearSystem.cpp:1096-1107) and v = w + dt/2 a (half_velocity).  w: the node's half-step velocity of
// the previous step (its velocity on constrained DOFs); rhs = -fint + fext of the node (read for free DOFs
// only).  With wkin, 1/2 m (V^2 - w^2) of each constrained DOF, whose velocity the prescribed V replaces, is
// added to *wkin.
__device__ __forceinline__ void node_update(const MpsInst& I, uint32_t info, long long k, double dt, double m,
                                            const double* rhs, const double* w, double* v, double* a,
                                            double* wkin = nullptr) {
#pragma clang fp contract(off)
    const uint32_t mask = info & 7u;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        if ((mask >> i) & 1u) {
            double U, V;
            bc_motion(I, I.bcval[3 * (size_t)((info & MPS_REC_MASK) >> 3) + i], k, dt, &U, &V);
            if (wkin) *wkin += 0.5 * m * (V * V - w[i] * w[i]);
            v[i] = V;
            a[i] = 0.;
        } else {
            const double an = (m != 0.0) ? rhs[i] / m : 0.0;
            v[i] = half_velocity(dt, w[i], an);
            a[i] = an;
        }
    }
}

// Gather + lumped-mass solve of node n at step k (node_partial_sum, node_fext, node_update): v, a of the step
// from the half-step velocity w.  A fully constrained node reads no forces.
__device__ __forceinline__ void solve_node(const MpsInst& I, int n, long long k, double dt, const double* w,
                                           double* v, double* a) {
    const uint32_t info = I.ninfo[n];
    double rhs[3] = {0.0, 0.0, 0.0};
    if ((info & 7u) != 7u) {
        double fi[3], fe[3];
        node_partial_sum(I, n, fi);
        node_fext(I, n, info, fe);
        node_rhs(fi, fe, rhs);
    }
    node_update(I, info, k, dt, I.mass[n], rhs, w, v, a);
}

// Energy bookkeeping of node n at step k, in two parts around its velocity update so that few values live
// through it.  Adds to red[MPS_NODE_PARTS]: the work of the external force over the step,
// 1/2 (fext_prev + fext) . incr; the work of the supports, 1/2 (r_prev + r) . incr + 1/2 m (v^2 - vold^2) over the
// constrained DOFs with r = fint - fext (the update sets a = 0 there, so the momentum the prescribed velocity
// gives the node's mass is the support's); the lumped kinetic energy 1/2 m v^2; and u . fext (EXTERNAL_ENERGY,
// dataExtraction.cpp:867-875).  Works are counted from step 2 on; every step stores the forces the next one
// takes as previous.
// Before the update: the force terms (u: displacement of the step, incr: its increment; fi, fe: internal and
// external force).
__device__ __forceinline__ void node_energy_forces(const MpsInst& I, int n, uint32_t info, long long k,
                                                   const double* u, const double* incr, const double* fi,
                                                   const double* fe, double* red) {
    const uint32_t mask = info & 7u;
    double fp[3] = {fe[0], fe[1], fe[2]};  // unflagged nodes: gravity alone (constant) or nothing
    if (info & MPS_NODE_LOADED) {
        double* lp = I.lfprev + 3 * (size_t)I.lidx[n];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            fp[i] = lp[i];
            lp[i] = fe[i];
        }
    }
    double we = -0.0, wr = -0.0, ee = -0.0;  // -0.0: the identity of the additions
#pragma unroll
    for (int i = 0; i < 3; i++) {
        we += 0.5 * (fp[i] + fe[i]) * incr[i];
        ee += u[i] * fe[i];
    }
    if (mask) {
        double* rp = I.rprev + 3 * (size_t)((info & MPS_REC_MASK) >> 3);
#pragma unroll
        for (int i = 0; i < 3; i++)
            if ((mask >> i) & 1u) {
                const double r = fi[i] - fe[i];
                wr += 0.5 * (rp[i] + r) * incr[i];
                rp[i] = r;
            }
    }
    if (k >= 2) {
        red[0] += we;
        red[1] += wr;
    }
    red[3] += ee;
}
// During the update (node_update's wkin): the kinetic term of the supports.  After it: the lumped kinetic
// energy (v: the new velocity).
__device__ __forceinline__ void node_energy_motion(double m, long long k, double wkin, const double* v, double* red) {
    double ek = -0.0;
#pragma unroll
    for (int i = 0; i < 3; i++) ek += 0.5 * m * v[i] * v[i];
    if (k >= 2) red[1] += wkin;
    red[2] += ek;
}

// ---------------------------------------------------------------------------
// XCD-aware tile mapping.  Workgroups are dispatched round-robin over the 8 XCDs
// (each with its own L2).  Remapping physical block b to logical tile
// L = x*q + min(x,r) + b/8 (x = b%8, q = nb/8, r = nb%8) gives every XCD one contiguous
// range of tiles, so mesh rows and layers that share nodes (element kernel) or element
// forces (node kernel) are processed by the same L2.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int xcd_tile(int b, int nb) {
    const int x = b & 7, i = b >> 3, q = nb >> 3, r = nb & 7;
    return x * q + (x < r ? x : r) + i;
}

// ---------------------------------------------------------------------------
// Reductions
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long umin64(unsigned long long a, unsigned long long b) { return b < a ? b : a; }

// Stable-dt reduction slots (MpsInst::dtred): workgroups atomicMin into one of MPS_DT_SLOTS slots of the
// step parity, 256 B apart, so that thousands of workgroups do not queue on one address.
__device__ __forceinline__ unsigned long long* dt_slot(const MpsInst& I, int parity, int s) {
    return I.dtred + (size_t)(parity * MPS_DT_SLOTS + s) * MPS_DT_STRIDE;
}
// Certificate counters (single-instance form): s < MPS_DT_SLOTS counts the element-kernel workgroups b with
// b % MPS_DT_SLOTS == s that are done, s = MPS_DT_SLOTS the residues whose workgroups are all done.
__device__ __forceinline__ unsigned int* dt_count(const MpsInst& I, int s) {
    return reinterpret_cast<unsigned int*>(I.dtred + (size_t)(2 * MPS_DT_SLOTS + s) * MPS_DT_STRIDE);
}
// Min over the slots of one parity, returned in every lane; all 64 lanes of the wave must call it.
__device__ __forceinline__ unsigned long long dt_read_min(const MpsInst& I, int parity) {
    unsigned long long v = *dt_slot(I, parity, threadIdx.x & (MPS_DT_SLOTS - 1));
#pragma unroll
    for (int off = MPS_DT_SLOTS / 2; off > 0; off >>= 1)
        v = umin64(v, (unsigned long long)__shfl_xor((long long)v, off, 64));
    return v;
}
// dt_read_min for a reader in the kernel whose workgroups are still adding to the slots (device-scope loads).
__device__ __forceinline__ unsigned long long dt_read_min_now(const MpsInst& I, int parity) {
    unsigned long long v = __hip_atomic_load(dt_slot(I, parity, threadIdx.x & (MPS_DT_SLOTS - 1)), __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
    for (int off = MPS_DT_SLOTS / 2; off > 0; off >>= 1)
        v = umin64(v, (unsigned long long)__shfl_xor((long long)v, off, 64));
    return v;
}
__device__ __forceinline__ void dt_reset(const MpsInst& I, int parity) {
#pragma unroll
    for (int s = 0; s < MPS_DT_SLOTS; s++) *dt_slot(I, parity, s) = 0x7FF0000000000000ull;
}

// Block-wide min of 64-bit patterns; result valid in thread 0.
__device__ __forceinline__ unsigned long long block_min_u64(unsigned long long v, unsigned long long* sh) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = umin64(v, (unsigned long long)__shfl_xor((long long)v, off, 64));
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        for (int w = 1; w < nw; w++) v = umin64(v, sh[w]);
    }
    return v;
}

// Deterministic block sum of NV doubles, result in thread 0: a fixed butterfly within each wave, then thread 0
// adds the waves' sums in wave order.  One barrier; sh holds NV doubles per wave.  A caller that reuses sh
// right after must synchronise first (thread 0 may still be reading it).
template <int NV>
__device__ __forceinline__ void block_sum(double* vals, double* sh) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int j = 0; j < NV; j++) vals[j] += __shfl_xor(vals[j], off, 64);
    const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int j = 0; j < NV; j++) sh[w * NV + j] = vals[j];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < NV; j++) {
            double t = sh[j];
            for (int q = 1; q < nw; q++) t += sh[q * NV + j];
            vals[j] = t;
        }
    }
}

// Grid-wide sums without fences or atomics: every block writes its block sums to part[blockIdx.x*NV + j]
// (publish_block_sum); a later kernel on the same stream, whose start makes them visible on every XCD, adds
// them in block order with one block (reduce_parts).  Deterministic, and no per-block device-scope release
// (on gfx950 a release writes the whole L2 back, which cost more than the rest of the energy work).
template <int NV>
__device__ __forceinline__ void publish_block_sum(double* vals, double* __restrict__ part, double* sh) {
    block_sum<NV>(vals, sh);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < NV; j++) part[(size_t)blockIdx.x * NV + j] = vals[j];
    }
}

// Sum of the nb block sums part[b*NV + j] by the calling block (every thread calls it): each thread adds the
// blocks b = threadIdx.x + i*blockDim.x in order (loaded 8 at a time, so that 8 loads are in flight), then
// block_sum.  Totals in out of thread 0.
template <int NV>
__device__ __forceinline__ void reduce_parts(const double* __restrict__ part, int nb, double* out, double* sh) {
    // 4 block sums per work-item in flight: more would set the register count of the whole kernel (the node
    // kernel of an energy step calls this from one block); the order of the additions does not depend on it
    constexpr int UN = 4;
    double acc[NV];
#pragma unroll
    for (int j = 0; j < NV; j++) acc[j] = 0.0;
    for (int b0 = threadIdx.x; b0 < nb; b0 += UN * blockDim.x) {
        double v[UN][NV];
#pragma unroll
        for (int q = 0; q < UN; q++) {
            const int b = b0 + q * blockDim.x;
#pragma unroll
            for (int j = 0; j < NV; j++) v[q][j] = b < nb ? part[(size_t)b * NV + j] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < UN; q++)
#pragma unroll
            for (int j = 0; j < NV; j++) acc[j] += v[q][j];
    }
    __syncthreads();  // sh may still be read by an earlier block_sum
    block_sum<NV>(acc, sh);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < NV; j++) out[j] = acc[j];
    }
}

__device__ __forceinline__ double* hist_at(const MpsInst& I, int col, long long k) {
    return I.hist + (size_t)col * I.hist_cap + (size_t)((k - 1) % I.hist_cap);
}

}  // namespace mpsd

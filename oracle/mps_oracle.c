/*
 * mps_oracle.c -- CPU restatement of the reference explicit solver.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, the smoke check and
 * the benchmark's cpu_baseline leg may load this library; the product path
 * (libmps_gpu.so) never links or calls it.
 *
 * Built with -O3 -ffp-contract=off (no FMA contraction), mirroring the
 * reference's x86-64 build flags (Makefile:18,34 of the reference: no -march,
 * no -ffast-math).  Each function restates one reference function and keeps
 * its floating-point operation order; the citation names the file:line.
 *
 * Two summation orders for the internal force are offered:
 *   ORC_SUM_REFERENCE  per-Gauss-point scatter into fint in GP order, exactly
 *                      as solverExplicit.cpp:363-379 (sequential only);
 *   ORC_SUM_ELEMENT    per-element sums over the element's GPs followed by a
 *                      node-centric gather in ascending (element, slot) order
 *                      (the association the GPU path uses).  For tet4 (one
 *                      GP per element) both orders are bitwise identical.
 */
#define _GNU_SOURCE
#include "mps_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* 3x3 algebra                                                         */
/* ------------------------------------------------------------------ */

/* determinantTensor, 3x3 branch: maths.cpp:1564-1577 */
double orc_det3(const double *A)
{
    return A[0] * (A[4] * A[8] - A[7] * A[5]) -
           A[1] * (A[3] * A[8] - A[6] * A[5]) +
           A[2] * (A[3] * A[7] - A[6] * A[4]);
}

/* inverse (adjugate / det), 3x3 branch: maths.cpp:2749-2767 */
void orc_inverse3(const double *A, double *Ainv)
{
    double det = orc_det3(A);
    double idet = 1.0 / det;
    Ainv[0] = idet * (A[4] * A[8] - A[5] * A[7]);
    Ainv[3] = -idet * (A[3] * A[8] - A[5] * A[6]);
    Ainv[6] = idet * (A[3] * A[7] - A[4] * A[6]);
    Ainv[1] = -idet * (A[1] * A[8] - A[2] * A[7]);
    Ainv[4] = idet * (A[0] * A[8] - A[2] * A[6]);
    Ainv[7] = -idet * (A[0] * A[7] - A[1] * A[6]);
    Ainv[2] = idet * (A[1] * A[5] - A[2] * A[4]);
    Ainv[5] = -idet * (A[0] * A[5] - A[2] * A[3]);
    Ainv[8] = idet * (A[0] * A[4] - A[1] * A[3]);
}

/* multGeneralMatrices: C[m x ll] = A[m x kk] * B[kk x ll], k innermost,
 * accumulating from zero: maths.cpp:1679-1693 */
static void mat_general(const double *A, int m, int kk, const double *B, int ll, double *C)
{
    for (int i = 0; i < m * ll; i++) C[i] = 0.0;
    for (int i = 0; i < m; i++)
        for (int j = 0; j < ll; j++)
            for (int k = 0; k < kk; k++)
                C[i * ll + j] += A[i * kk + k] * B[k * ll + j];
}

/* ------------------------------------------------------------------ */
/* Shape functions and quadrature                                      */
/* ------------------------------------------------------------------ */

/* classLinearHexahedra::getInterpolationParameters: classLinearHexahedra.cpp:344-381 */
void orc_hex8_shape(const double *xw, double *phi, double *nd)
{
    const double x = xw[0], y = xw[1], z = xw[2];
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

/* classLinearTetrahedra::getInterpolationParameters: classLinearTetrahedra.cpp:270-292 */
void orc_tet4_shape(const double *xw, double *phi, double *nd)
{
    phi[0] = -1.0 * xw[0] - 1.0 * xw[1] - 1.0 * xw[2] + 1.0;
    phi[1] = 1.0 * xw[0];
    phi[2] = 1.0 * xw[1];
    phi[3] = 1.0 * xw[2];
    nd[0] = -1.0 - 0.0 - 0.0 + 0.0;
    nd[3] = 1.0;
    nd[6] = 0.0;
    nd[9] = 0.0;
    nd[1] = -0.0 - 1.0 - 0.0 + 0.0;
    nd[4] = 0.0;
    nd[7] = 1.0;
    nd[10] = 0.0;
    nd[2] = -0.0 - 0.0 - 1.0 + 0.0;
    nd[5] = 0.0;
    nd[8] = 0.0;
    nd[11] = 1.0;
}

/* GPsHexahedron3DValues(order 2) / GPsTetrahedron3DValues(order 1):
 * GPsDistribution.cpp:392-432 and 339-343.  xw is [ngp][4] (xi,eta,zeta,w). */
int32_t orc_quadrature(int32_t nen, double *xw)
{
    if (nen == 8) {
        static const double s[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                       {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
        for (int g = 0; g < 8; g++) {
            for (int d = 0; d < 3; d++) xw[g * 4 + d] = s[g][d] > 0 ? 0.577350269 : -0.577350269;
            xw[g * 4 + 3] = 1;
        }
        return 8;
    }
    if (nen == 4) {
        xw[0] = 0.250000000000;
        xw[1] = 0.25;
        xw[2] = 0.25;
        xw[3] = 0.166666666667;
        return 1;
    }
    return 0;
}

/* One Gauss point of classLinearHexahedra::defineGPs (classLinearHexahedra.cpp:242-342)
 * or classLinearTetrahedra::defineGPs (classLinearTetrahedra.cpp:171-268), plus the
 * classGPs constructor (classGPs.cpp:24-31): wJ = weight*|det J|.
 * xyz is node-major [a*3+d]; weight is 1*xw[3] for hex, xw[3] for tet. */
void orc_define_gp(int32_t nen, const double *xyz, const double *xw,
                   double *phi, double *dphi, double *wJ, double *detJ)
{
    double xyzCoor[24], nd[24], Jac[9], invJ[9];
    for (int i = 0; i < nen; i++) {
        xyzCoor[i] = xyz[i * 3 + 0];
        xyzCoor[nen + i] = xyz[i * 3 + 1];
        xyzCoor[2 * nen + i] = xyz[i * 3 + 2];
    }
    if (nen == 8) orc_hex8_shape(xw, phi, nd);
    else orc_tet4_shape(xw, phi, nd);
    mat_general(xyzCoor, 3, nen, nd, 3, Jac);
    double J = orc_det3(Jac);
    orc_inverse3(Jac, invJ);
    double weight = (nen == 8) ? 1 * xw[3] : xw[3];
    mat_general(nd, nen, 3, invJ, 3, dphi);
    *wJ = weight * fabs(J);
    if (detJ) *detJ = fabs(J);
}

/* ------------------------------------------------------------------ */
/* Characteristic length (stable time step)                            */
/* ------------------------------------------------------------------ */

/* areaTriangle, 3D branch: geometry.cpp:112-130.  tri = {x1,x2,x3,y1,y2,y3,z1,z2,z3} */
double orc_area_triangle(const double *t)
{
    double A[3], B[3], C[3];
    for (int i = 0; i < 3; i++) {
        A[i] = t[i * 3] - t[i * 3 + 2];
        B[i] = t[i * 3 + 1] - t[i * 3 + 2];
    }
    /* crossProduct: geometry.cpp:180-190 */
    C[0] = A[1] * B[2] - A[2] * B[1];
    C[1] = A[2] * B[0] - A[0] * B[2];
    C[2] = A[0] * B[1] - A[1] * B[0];
    return 0.5 * sqrt(C[0] * C[0] + C[1] * C[1] + C[2] * C[2]);
}

/* areaSquare: triangles (0,1,2)+(0,2,3): geometry.cpp:143-174.
 * quad = {x0..x3, y0..y3, z0..z3} */
static double area_square(const double *q)
{
    double t1[9], t2[9];
    for (int i = 0; i < 3; i++) {
        t1[i * 3] = q[i * 4];
        t1[i * 3 + 1] = q[i * 4 + 1];
        t1[i * 3 + 2] = q[i * 4 + 2];
        t2[i * 3] = q[i * 4];
        t2[i * 3 + 1] = q[i * 4 + 2];
        t2[i * 3 + 2] = q[i * 4 + 3];
    }
    double a1 = orc_area_triangle(t1);
    double a2 = orc_area_triangle(t2);
    return a1 + a2;
}

/* inscribedSphereHex: geometry.cpp:308-355 (V is the corner parallelepiped) */
static double inscribed_sphere_hex(const double *p /* node-major [8][3] */)
{
    static const int faces[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 1, 5, 4},
                                    {1, 2, 6, 5}, {2, 3, 7, 6}, {0, 3, 7, 4}};
    double s[6];
    for (int f = 0; f < 6; f++) {
        double q[12];
        for (int d = 0; d < 3; d++)
            for (int k = 0; k < 4; k++) q[d * 4 + k] = p[faces[f][k] * 3 + d];
        s[f] = area_square(q);
    }
    double A = fabs(s[0]) + fabs(s[1]) + fabs(s[2]) + fabs(s[3]) + fabs(s[4]) + fabs(s[5]);
    double M[9] = {p[3] - p[0], p[4] - p[1], p[5] - p[2],
                   p[9] - p[0], p[10] - p[1], p[11] - p[2],
                   p[12] - p[0], p[13] - p[1], p[14] - p[2]};
    double V = fabs(orc_det3(M));
    return 3.0 * V / A;
}

/* inscribedSphere (tetrahedron): geometry.cpp:264-301 */
static double inscribed_sphere_tet(const double *p /* node-major [4][3] */)
{
    static const int tris[4][3] = {{0, 1, 2}, {1, 2, 3}, {0, 2, 3}, {0, 1, 3}};
    double s[4];
    for (int f = 0; f < 4; f++) {
        double t[9];
        for (int d = 0; d < 3; d++)
            for (int k = 0; k < 3; k++) t[d * 3 + k] = p[tris[f][k] * 3 + d];
        s[f] = orc_area_triangle(t);
    }
    double A = fabs(s[0]) + fabs(s[1]) + fabs(s[2]) + fabs(s[3]);
    double M[9] = {p[3] - p[0], p[4] - p[1], p[5] - p[2],
                   p[6] - p[0], p[7] - p[1], p[8] - p[2],
                   p[9] - p[0], p[10] - p[1], p[11] - p[2]};
    double V = 1.0 / 6.0 * fabs(orc_det3(M));
    return 3.0 * V / A;
}

/* characteristicLength = 2*inscribed radius:
 * classLinearHexahedra.cpp:209-234, classLinearTetrahedra.cpp:146-163 */
double orc_char_length(int32_t nen, const double *xyz)
{
    double r = (nen == 8) ? inscribed_sphere_hex(xyz) : inscribed_sphere_tet(xyz);
    return 2 * r;
}

/* ------------------------------------------------------------------ */
/* Constitutive laws                                                   */
/* ------------------------------------------------------------------ */

/* Material constants: hyperElasticNeoHookean.cpp:28-30 and 180-181 */
void orc_material(double E, double nu, double *lambda, double *mu, double *K1)
{
    *lambda = (E * nu) / ((1. + nu) * (1. - 2. * nu));
    *mu = 0.5 * E / (1. + nu);
    *K1 = E / (3 * (1 - 2 * nu));
}

/* soundSpeed: hyperElasticNeoHookean.cpp:51-54 (identical at 202-205) */
double orc_sound_speed(double E, double nu, double rho)
{
    double lambda, mu, K1;
    orc_material(E, nu, &lambda, &mu, &K1);
    return sqrt((lambda + 2 * mu) / rho);
}

/* classHyperElastic3DCompressibleNeoHookean::updateConstitutive (no tangent):
 * hyperElasticNeoHookean.cpp:221-257 with TensorOperations::determinant
 * (TensorOperations.h:56-74), rightCauchy (:478-492), inverseTensor2 (:22-26),
 * classTensor2::addDiagonal/axpy (maths.h:350-366) and dot (:215-229). */
static void pk1_compressible(const double *F, double lambda, double mu, double *P)
{
    double J = orc_det3(F);
    double C[9], invC[9], S[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double c = 0;
            for (int k = 0; k < 3; k++) c += F[k * 3 + i] * F[k * 3 + j];
            C[i * 3 + j] = c;
        }
    orc_inverse3(C, invC);
    for (int i = 0; i < 9; i++) S[i] = 0;
    for (int i = 0; i < 3; i++) S[i * 3 + i] += mu;
    double alpha = lambda * log(J) - mu;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) S[i * 3 + j] += alpha * invC[i * 3 + j];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double p = 0;
            for (int k = 0; k < 3; k++) p += F[i * 3 + k] * S[k * 3 + j] * 1.;
            P[i * 3 + j] = p;
        }
}

/* classHyperElasticNeoHookean::updateConstitutive (no tangent):
 * hyperElasticNeoHookean.cpp:70-126 with leftCauchy (TensorOperations.h:517-531),
 * classTensor2::trace (maths.h:303-311) and dotSecondTranspose (:577-591). */
static void pk1_belytschko_sig(const double *F, double mu, double K1, double *P, double *sig_out)
{
    double B[9], invF[9], sig[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double b = 0;
            for (int k = 0; k < 3; k++) b += F[i * 3 + k] * F[j * 3 + k];
            B[i * 3 + j] = b;
        }
    double J = orc_det3(F);
    orc_inverse3(F, invF);
    double trB = 0;
    for (int i = 0; i < 3; i++) trB += B[i * 3 + i];
    for (int i = 0; i < 9; i++) sig[i] = 0;
    double fact1 = pow(J, 5.0 / (double)3);
    double fact2 = mu / fact1;
    for (int i = 0; i < 9; i++) sig[i] += fact2 * B[i];
    double fact3 = -(mu / (double)3) * trB / fact1 + K1 * (J - 1.0);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) sig[i * 3 + j] += fact3 * (i == j ? 1.0 : 0.0);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double p = 0;
            for (int k = 0; k < 3; k++) p += sig[i * 3 + k] * invF[j * 3 + k] * J;
            P[i * 3 + j] = p;
        }
    if (sig_out)
        for (int i = 0; i < 9; i++) sig_out[i] = sig[i];
}

static void pk1_belytschko(const double *F, double mu, double K1, double *P)
{
    pk1_belytschko_sig(F, mu, K1, P, NULL);
}

/* PK2 = mu2 E + lam tr(E) I (hyperElasticStVenantKirchhoff.cpp / linearElastic.cpp stress(), 3-D branch):
 * PK2 = E; multiSTensorScalar(PK2, mu2); PK2_ii += lam*traceE.  mu2 = E/(1+nu) equals 2*mu bitwise
 * (mu = 0.5*E/(1+nu) differs by an exact power-of-two scaling); lam is the NH lambda expression. */
static void stress_lin(const double *E, double lambda, double mu, double *S)
{
    double mu2 = 2.0 * mu, trE = 0;
    for (int i = 0; i < 3; i++) trE += E[i * 3 + i];
    for (int i = 0; i < 9; i++) S[i] = E[i] * mu2;
    for (int i = 0; i < 3; i++) S[i * 3 + i] += lambda * trE;
}

/* Green-Lagrange strain: computeGL (maths.cpp:1227-1240) with multSTensor3FirstTranspose
 * (maths.cpp:1760-1772, accumulating into a zeroed C). */
static void green_lagrange(const double *F, double *E)
{
    double C[9] = {0};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) C[i * 3 + j] += F[k * 3 + i] * F[k * 3 + j];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) E[i * 3 + j] = 0.5 * (C[i * 3 + j] - (i == j ? 1.0 : 0.0));
}

/* classHyperElasticStVenantKirchhoff::updateConstitutive (no tangent): hyperElasticStVenantKirchhoff.cpp:203-226;
 * PK1 = F PK2 by multTensorTensor3 (maths.cpp:1703-1715) into the PK1 zeroed by startFromOtherState. */
static void pk1_svk(const double *F, double lambda, double mu, double *P)
{
    double E[9], S[9];
    green_lagrange(F, E);
    stress_lin(E, lambda, mu, S);
    for (int i = 0; i < 9; i++) P[i] = 0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) P[i * 3 + j] += F[i * 3 + k] * S[k * 3 + j];
}

/* linearElastic::updateConstitutive: linearElastic.cpp:63-95, E = sym(F) - I, PK1 = stress(E). */
static void pk1_linear(const double *F, double lambda, double mu, double *P)
{
    double E[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) E[i * 3 + j] = 0.5 * (F[i * 3 + j] + F[j * 3 + i] - 2. * (i == j ? 1.0 : 0.0));
    stress_lin(E, lambda, mu, P);
}

static void pk1_law(int32_t law, const double *F, double lambda, double mu, double K1, double *P)
{
    if (law == ORC_LAW_NH_BELYTSCHKO) pk1_belytschko(F, mu, K1, P);
    else if (law == ORC_LAW_SVK) pk1_svk(F, lambda, mu, P);
    else if (law == ORC_LAW_LINEAR) pk1_linear(F, lambda, mu, P);
    else pk1_compressible(F, lambda, mu, P);
}

void orc_pk1(int32_t law, const double *F, double E, double nu, double *P)
{
    double lambda, mu, K1;
    orc_material(E, nu, &lambda, &mu, &K1);
    pk1_law(law, F, lambda, mu, K1, P);
}

/* soundSpeed: NH hyperElasticNeoHookean.cpp:51-54, 202-205; St-Venant-Kirchhoff
 * hyperElasticStVenantKirchhoff.cpp (sqrt(E/rho)); linearElastic.cpp:49-54. */
double orc_law_sound_speed(int32_t law, double E, double nu, double rho)
{
    if (law == ORC_LAW_SVK || law == ORC_LAW_LINEAR) return sqrt(E / rho);
    return orc_sound_speed(E, nu, rho);
}

/* classHyperElastic3DCompressibleNeoHookean::defoEnergy: hyperElasticNeoHookean.cpp:291-304.
 * The Belytschko law does not override defoEnergy and returns 0 (constitutiveModels.h:91-94). */
static double defo_energy_lm(int32_t law, const double *F, double lambda, double mu)
{
    if (law == ORC_LAW_NH_BELYTSCHKO) return 0.0;
    if (law == ORC_LAW_SVK) {  /* hyperElasticStVenantKirchhoff.cpp:251-274: 0.5 PK2:E */
        double E[9], S[9], w = 0;
        green_lagrange(F, E);
        stress_lin(E, lambda, mu, S);
        for (int i = 0; i < 9; i++) w += S[i] * E[i];
        return 0.5 * w;
    }
    if (law == ORC_LAW_LINEAR) {  /* linearElastic.cpp:97-110: 0.5 (P:F - P:I) */
        double P[9], pf = 0, pd = 0;
        pk1_linear(F, lambda, mu, P);
        for (int i = 0; i < 9; i++) pf += P[i] * F[i];
        for (int i = 0; i < 3; i++) pd += P[i * 3 + i];
        return 0.5 * (pf - pd);
    }
    double J = orc_det3(F);
    double trC = 0;
    for (int i = 0; i < 3; i++) {
        double c = 0;
        for (int k = 0; k < 3; k++) c += F[k * 3 + i] * F[k * 3 + i];
        trC += c;
    }
    return 0.5 * lambda * log(J) * log(J) - mu * log(J) + 0.5 * mu * trC;
}

double orc_defo_energy(int32_t law, const double *F, double E, double nu)
{
    double lambda, mu, K1;
    orc_material(E, nu, &lambda, &mu, &K1);
    return defo_energy_lm(law, F, lambda, mu);
}

/* ------------------------------------------------------------------ */
/* Per Gauss point kinematics, force and mass                          */
/* ------------------------------------------------------------------ */

/* constitutiveManager::computeKinematicVariables: constitutiveManager.cpp:150-164.
 * uloc is [a*3+i] (node a, direction i). */
void orc_kinematics(int32_t nen, const double *dphi, const double *uloc, double *F)
{
    for (int i = 0; i < 9; i++) F[i] = 0.;
    for (int i = 0; i < 3; i++) {
        F[i * 3 + i] += 1.;
        for (int a = 0; a < nen; a++)
            for (int j = 0; j < 3; j++) F[i * 3 + j] += dphi[a * 3 + j] * uloc[a * 3 + i];
    }
}

/* MechanicalTerm::getForceVector: constitutiveAlgorithms.cpp:33-57.
 * f is indexed like vForce: f[a + i*nen]. */
void orc_gp_force(int32_t nen, double wJ, const double *P, const double *dphi, double *f)
{
    for (int k = 0; k < 3 * nen; k++) f[k] = 0;
    for (int a = 0; a < nen; a++)
        for (int i = 0; i < 3; i++)
            for (int J = 0; J < 3; J++) f[a + i * nen] += wJ * P[i * 3 + J] * dphi[a * 3 + J];
}

/* MechanicalTerm::getLumpedMassVector, linear-element branch:
 * constitutiveAlgorithms.cpp:191-205.  m[a] (identical for the three directions). */
void orc_gp_lumped_mass(int32_t nen, double wJ, double rho, const double *phi, double *m)
{
    for (int a = 0; a < nen; a++) {
        double s = 0;
        for (int b = 0; b < nen; b++) s += wJ * rho * phi[a] * phi[b];
        m[a] = s;
    }
}

/* ------------------------------------------------------------------ */
/* Solver state                                                        */
/* ------------------------------------------------------------------ */

/* Per Neumann block: what its constructor and the set-up fix (the rest is evaluated every step). */
struct load_state {
    double *fm;        /* FORCE: [3n] F/(tf-t0) (classNeumann ctor, classNodes.cpp:85-96) */
    double *tg;        /* GRAVITY: [Ne*ngp] -w rho |J| g per bulk GP */
    int ns;            /* PRESSURE: surface GPs (n faces x 4 or 1) */
    int32_t *se;       /* [ns] bulk element */
    int32_t *sn;       /* [ns*4] face nodes */
    double *ws;        /* [ns] w |J_s| */
    double *n0;        /* [ns*3] reference normal of the face */
    double *phs;       /* [ns*4] face shape values */
    double *phb;       /* [ns*8] bulk shape values at the point */
    double *dphb;      /* [ns*24] bulk gradients at the point */
};

typedef struct {
    const orc_problem *p;
    int ngp, nen, N, Ne;
    double *phi;       /* [Ne*ngp*nen] */
    double *dphi;      /* [Ne*ngp*nen*3] */
    double *wJ;        /* [Ne*ngp] */
    double *l0;        /* [Ne] initial characteristic length */
    double *c;         /* [Ne] sound speed */
    double *lam, *mu, *K1, *rho;
    double *mass;      /* [N] */
    int8_t *active;    /* [3N] */
    int32_t *rowptr, *col; /* CSR node -> (e*nen + a) ascending */
    double *fe;        /* [Ne*nen*3] element forces, [(e*nen+a)*3+i] */
    double *fint;      /* [3N] */
    double *fext;      /* [3N] external force of the current step */
    struct load_state *ld; /* [n_loads] */
    /* threading */
    int nt;
    double *tmin;      /* per-thread partials */
    double *tsum;      /* per-thread partials, 3 per thread */
} orc_state;

static double mat_at(const double *v, int per_elem, int e) { return per_elem ? v[e] : v[0]; }

static void gather_xyz(const orc_state *s, int e, const double *u, double *xyz)
{
    const int32_t *cn = s->p->conn + (size_t)e * s->nen;
    for (int a = 0; a < s->nen; a++)
        for (int d = 0; d < 3; d++) {
            double x = s->p->X[(size_t)d * s->N + cn[a]];
            if (u) x += u[(size_t)d * s->N + cn[a]];
            xyz[a * 3 + d] = x;
        }
}

/* Faces of getSubElement: hex (classLinearHexahedra.cpp:30-85), tet (classLinearTetrahedra.cpp:30-55). */
static const int HEX_FACE[6][4] = {{0, 1, 2, 3}, {4, 7, 6, 5}, {0, 4, 5, 1}, {1, 5, 6, 2}, {2, 6, 7, 3}, {3, 7, 4, 0}};
static const int TET_FACE[4][3] = {{0, 1, 2}, {0, 3, 1}, {1, 3, 2}, {0, 2, 3}};

/* norm2: maths.cpp:1550-1557 */
static double norm2_3(const double *a)
{
    double t = 0;
    for (int i = 0; i < 3; i++) t += a[i] * a[i];
    return sqrt(t);
}

/* crossProduct: geometry.cpp:180-190 */
static void cross3(const double *A, const double *B, double *C)
{
    C[0] = A[1] * B[2] - A[2] * B[1];
    C[1] = A[2] * B[0] - A[0] * B[2];
    C[2] = A[0] * B[1] - A[1] * B[0];
}

/* Surface GPs of face S of bulk element e, each with the bulk shape at its position:
 *   face nodes and getn0 (classLinearSquare.cpp:19-45, classLinearTriangles.cpp:19-45: (p2-p1)x(p3-p1)/|.|);
 *   defineGPs of the face (classLinearSquare.cpp:196-301, 2x2 points +-0.577350269 weight 1;
 *   classLinearTriangles.cpp:194-299, one point (0.33333333333333, 0.33333333333333) weight 0.5*1),
 *   in the local plane basis of createOrthogonalBasisOnPlane (geometry.cpp:363-389): J = det of the 2x2
 *   Jacobian, position = sum_k x_k phi_k;
 *   defineShapeOnAPoint of the bulk element at that position (classLinearHexahedra.cpp:95-172: natural
 *   coordinates 2 A^-1 (x - x_0) - 1 with A the corner edges 0-1, 0-3, 0-4; classLinearTetrahedra.cpp:66-117:
 *   the shape at the natural origin, phi = (1, 0, 0, 0)).
 * Returns the number of points or ORC_ERR_ARG. */
static int surface_gps(const orc_state *s, int e, int S, int32_t *sn, double *pos, double *ws, double *n0,
                       double *phs, double *phb, double *dphb)
{
    const int nen = s->nen, nf = (nen == 8) ? 4 : 3;
    if (e < 0 || e >= s->Ne || S < 1 || S > ((nen == 8) ? 6 : 4)) return ORC_ERR_ARG;
    const int *fl = (nen == 8) ? HEX_FACE[S - 1] : TET_FACE[S - 1];
    double xb[24], xf[4][3];
    gather_xyz(s, e, NULL, xb);
    for (int k = 0; k < nf; k++) {
        sn[k] = s->p->conn[(size_t)e * nen + fl[k]];
        for (int d = 0; d < 3; d++) xf[k][d] = xb[fl[k] * 3 + d];
    }
    double nn[3];
    for (int i = 0; i < 3; i++) {
        const int c1 = (i == 0) ? 1 : (i == 1) ? 2 : 0, c2 = (i == 0) ? 2 : (i == 1) ? 0 : 1;
        nn[i] = (xf[1][c1] - xf[0][c1]) * (xf[2][c2] - xf[0][c2]) - (xf[2][c1] - xf[0][c1]) * (xf[1][c2] - xf[0][c2]);
    }
    const double no = norm2_3(nn);
    double a[3], b[3], normal[3], xhat[3], map[6];
    for (int i = 0; i < 3; i++) {
        a[i] = xf[1][i] - xf[0][i];
        b[i] = xf[2][i] - xf[0][i];
    }
    cross3(b, a, normal);
    cross3(a, normal, xhat);
    const double magA = norm2_3(a), magX = norm2_3(xhat);
    for (int i = 0; i < 3; i++) {
        map[3 + i] = a[i] / magA;
        map[i] = xhat[i] / magX;
    }
    double tmp[12], loc[8];
    for (int k = 0; k < nf; k++)
        for (int d = 0; d < 3; d++) tmp[d * nf + k] = xf[k][d] - xf[0][d];
    mat_general(map, 2, 3, tmp, nf, loc);
    const double G = 0.577350269;
    const double sq[4][2] = {{G, G}, {-G, G}, {G, -G}, {-G, -G}};
    const int ng = (nen == 8) ? 4 : 1;
    for (int g = 0; g < ng; g++) {
        double ph[4], nd[8], Jac[4], weight;
        if (nen == 8) {
            const double x = sq[g][0], y = sq[g][1];
            ph[0] = 0.25 * (1 - x) * (1 - y);
            ph[1] = 0.25 * (1 + x) * (1 - y);
            ph[2] = 0.25 * (1 + x) * (1 + y);
            ph[3] = 0.25 * (1 - x) * (1 + y);
            nd[0] = -0.25 * (1 - y);
            nd[2] = 0.25 * (1 - y);
            nd[4] = 0.25 * (1 + y);
            nd[6] = -0.25 * (1 + y);
            nd[1] = -0.25 * (1 - x);
            nd[3] = -0.25 * (1 + x);
            nd[5] = 0.25 * (1 + x);
            nd[7] = 0.25 * (1 - x);
            weight = 1 * 1.0;
        } else {
            const double x = 0.33333333333333, y = 0.33333333333333;
            ph[0] = -1.0 * x - 1.0 * y + 1.0;
            ph[1] = 1.0 * x;
            ph[2] = 1.0 * y;
            ph[3] = 0.0;
            nd[0] = -1.0 - 0.0 + 0.0;
            nd[2] = 1.0;
            nd[4] = 0.0;
            nd[1] = -0.0 - 1.0 + 0.0;
            nd[3] = 0.0;
            nd[5] = 1.0;
            weight = 0.5 * 1.0;
        }
        mat_general(loc, 2, nf, nd, 2, Jac);
        const double J = Jac[0] * Jac[3] - Jac[2] * Jac[1];
        double x[3];
        for (int i = 0; i < 3; i++) {
            x[i] = 0;
            for (int k = 0; k < nf; k++) x[i] += xf[k][i] * ph[k];
        }
        double xw[4] = {0, 0, 0, 1};
        if (nen == 8) {
            double A[9], Ai[9], X[3];
            const int c[3] = {1, 3, 4};
            for (int d = 0; d < 3; d++)
                for (int j = 0; j < 3; j++) A[d * 3 + j] = xb[c[j] * 3 + d] - xb[d];
            for (int d = 0; d < 3; d++) X[d] = x[d] - xb[d];
            orc_inverse3(A, Ai);
            for (int i = 0; i < 3; i++) {
                double r = 0;
                for (int j = 0; j < 3; j++) r += Ai[i * 3 + j] * X[j];
                xw[i] = 2 * r - 1;
            }
        }
        double wJb, dJb;
        orc_define_gp(nen, xb, xw, phb + g * 8, dphb + g * 24, &wJb, &dJb);
        ws[g] = weight * fabs(J);
        for (int i = 0; i < 3; i++) {
            pos[g * 3 + i] = x[i];
            n0[g * 3 + i] = nn[i] / no;
        }
        for (int k = 0; k < 4; k++) phs[g * 4 + k] = (k < nf) ? ph[k] : 0.0;
        if (g > 0)
            for (int k = 0; k < 4; k++) sn[g * 4 + k] = sn[k];
    }
    return ng;
}

/* Set-up of the Neumann blocks: the classNeumann constructor's F/(tf-t0), the gravity factors of every
 * bulk GP, and the surface GPs of every pressure face (built once here; the reference rebuilds them each
 * step from the same reference coordinates, boundaryConditions.cpp:374-376). */
static int loads_init(orc_state *s, const double *xw)
{
    const orc_problem *p = s->p;
    const size_t N = s->N, Ne = s->Ne, ng = s->ngp;
    s->fext = calloc(3 * N, sizeof(double));
    if (!s->fext) return ORC_ERR_MEM;
    if (p->n_loads <= 0) return ORC_OK;
    if (!p->loads) return ORC_ERR_ARG;
    s->ld = calloc((size_t)p->n_loads, sizeof(struct load_state));
    if (!s->ld) return ORC_ERR_MEM;
    for (int b = 0; b < p->n_loads; b++) {
        const orc_load *L = &p->loads[b];
        struct load_state *l = &s->ld[b];
        if (L->n < 0) return ORC_ERR_ARG;
        if (L->type == ORC_LOAD_FORCE) {
            l->fm = malloc(sizeof(double) * 3 * ((size_t)L->n + 1));
            if (!l->fm) return ORC_ERR_MEM;
            for (int k = 0; k < L->n; k++) {
                if (L->index[k] < 0 || L->index[k] >= s->N) return ORC_ERR_ARG;
                for (int d = 0; d < 3; d++) l->fm[3 * k + d] = L->force[3 * k + d] / (L->tf - L->t0);
            }
        } else if (L->type == ORC_LOAD_GRAVITY) {  /* -w * rho * |J| * g, times phi_a per node */
            l->tg = malloc(sizeof(double) * Ne * ng);
            if (!l->tg) return ORC_ERR_MEM;
            for (size_t e = 0; e < Ne; e++) {
                double xyz[24];
                gather_xyz(s, (int)e, NULL, xyz);
                for (size_t g = 0; g < ng; g++) {
                    double phi[8], dphi[24], wJ, detJ;
                    orc_define_gp(p->nen, xyz, xw + g * 4, phi, dphi, &wJ, &detJ);
                    const double w = xw[g * 4 + 3], J = fabs(detJ);
                    l->tg[e * ng + g] = -w * s->rho[e] * J * L->value;
                }
            }
        } else if (L->type >= ORC_LOAD_PRESSURE_RAMP && L->type <= ORC_LOAD_TRACTION_INST) {
            if (L->type >= ORC_LOAD_TRACTION_RAMP && !L->force) return ORC_ERR_ARG;
            const int per = (s->nen == 8) ? 4 : 1;
            const size_t ns = (size_t)L->n * per + 1;
            l->se = malloc(sizeof(int32_t) * ns);
            l->sn = malloc(sizeof(int32_t) * ns * 4);
            l->ws = malloc(sizeof(double) * ns);
            l->n0 = malloc(sizeof(double) * ns * 3);
            l->phs = malloc(sizeof(double) * ns * 4);
            l->phb = malloc(sizeof(double) * ns * 8);
            l->dphb = malloc(sizeof(double) * ns * 24);
            if (!l->se || !l->sn || !l->ws || !l->n0 || !l->phs || !l->phb || !l->dphb) return ORC_ERR_MEM;
            for (int f = 0; f < L->n; f++) {
                double pos[12];
                const int g0 = l->ns;
                const int r = surface_gps(s, L->index[f], L->face[f], l->sn + 4 * g0, pos, l->ws + g0, l->n0 + 3 * g0,
                                          l->phs + 4 * g0, l->phb + 8 * g0, l->dphb + 24 * g0);
                if (r < 0) return r;
                for (int g = 0; g < r; g++) l->se[g0 + g] = L->index[f];
                l->ns += r;
            }
        } else {
            return ORC_ERR_ARG;
        }
    }
    return ORC_OK;
}

static int state_init(orc_state *s, const orc_problem *p)
{
    memset(s, 0, sizeof(*s));
    s->p = p;
    s->nen = p->nen;
    s->N = p->n_nodes;
    s->Ne = p->n_elem;
    double xw[32];
    s->ngp = orc_quadrature(p->nen, xw);
    if (s->ngp == 0) return ORC_ERR_ARG;
    size_t Ne = (size_t)s->Ne, ng = (size_t)s->ngp, nen = (size_t)s->nen, N = (size_t)s->N;
    s->phi = malloc(sizeof(double) * Ne * ng * nen);
    s->dphi = malloc(sizeof(double) * Ne * ng * nen * 3);
    s->wJ = malloc(sizeof(double) * Ne * ng);
    s->l0 = malloc(sizeof(double) * Ne);
    s->c = malloc(sizeof(double) * Ne);
    s->lam = malloc(sizeof(double) * Ne);
    s->mu = malloc(sizeof(double) * Ne);
    s->K1 = malloc(sizeof(double) * Ne);
    s->rho = malloc(sizeof(double) * Ne);
    s->mass = calloc(N, sizeof(double));
    s->active = malloc(3 * N);
    s->rowptr = calloc(N + 1, sizeof(int32_t));
    s->col = malloc(sizeof(int32_t) * Ne * nen);
    s->fe = malloc(sizeof(double) * Ne * nen * 3);
    s->fint = malloc(sizeof(double) * 3 * N);
    if (!s->phi || !s->dphi || !s->wJ || !s->l0 || !s->c || !s->lam || !s->mu || !s->K1 ||
        !s->rho || !s->mass || !s->active || !s->rowptr || !s->col || !s->fe || !s->fint)
        return ORC_ERR_MEM;
    /* GP geometry (classLinear*::defineGPs), initial char length (element ctor) */
    for (size_t e = 0; e < Ne; e++) {
        double xyz[24];
        gather_xyz(s, (int)e, NULL, xyz);
        for (size_t g = 0; g < ng; g++) {
            size_t gi = e * ng + g;
            orc_define_gp(p->nen, xyz, xw + g * 4, s->phi + gi * nen, s->dphi + gi * nen * 3,
                          s->wJ + gi, NULL);
        }
        s->l0[e] = orc_char_length(p->nen, xyz);
        double E = mat_at(p->E, p->mat_per_elem, (int)e), nu = mat_at(p->nu, p->mat_per_elem, (int)e);
        s->rho[e] = mat_at(p->rho, p->mat_per_elem, (int)e);
        orc_material(E, nu, s->lam + e, s->mu + e, s->K1 + e);
        s->c[e] = (p->law == ORC_LAW_SVK || p->law == ORC_LAW_LINEAR) ? sqrt(E / s->rho[e])
                                                                      : sqrt((s->lam[e] + 2 * s->mu[e]) / s->rho[e]);
    }
    /* active DOF map (dispBCManagement, inputFile.cpp:3185-3212) */
    memset(s->active, 1, 3 * N);
    for (int k = 0; k < p->n_dir; k++) {
        int32_t d = p->dir_dof[k];
        if (d < 0 || (size_t)d >= 3 * N) return ORC_ERR_ARG;
        s->active[d] = 0;
    }
    /* CSR node -> (element, slot), ascending element then slot */
    for (size_t e = 0; e < Ne; e++)
        for (size_t a = 0; a < nen; a++) {
            int32_t n = p->conn[e * nen + a];
            if (n < 0 || n >= s->N) return ORC_ERR_ARG;
            s->rowptr[n + 1]++;
        }
    for (size_t n = 0; n < N; n++) s->rowptr[n + 1] += s->rowptr[n];
    int32_t *fill = malloc(sizeof(int32_t) * N);
    if (!fill) return ORC_ERR_MEM;
    memcpy(fill, s->rowptr, sizeof(int32_t) * N);
    for (size_t e = 0; e < Ne; e++)
        for (size_t a = 0; a < nen; a++) {
            int32_t n = p->conn[e * nen + a];
            s->col[fill[n]++] = (int32_t)(e * nen + a);
        }
    free(fill);
    return loads_init(s, xw);
}

static void state_free(orc_state *s)
{
    if (s->ld)
        for (int k = 0; k < s->p->n_loads; k++) {
            struct load_state *l = &s->ld[k];
            free(l->fm); free(l->tg); free(l->se); free(l->sn); free(l->ws); free(l->n0);
            free(l->phs); free(l->phb); free(l->dphb);
        }
    free(s->ld);
    free(s->phi); free(s->dphi); free(s->wJ); free(s->l0); free(s->c);
    free(s->lam); free(s->mu); free(s->K1); free(s->rho); free(s->mass);
    free(s->active); free(s->rowptr); free(s->col); free(s->fe); free(s->fint);
    free(s->tmin); free(s->tsum); free(s->fext);
}

/* ------------------------------------------------------------------ */
/* Minimal thread pool (deterministic static partition)                */
/* ------------------------------------------------------------------ */

typedef void (*task_fn)(orc_state *s, int tid, int nt, void *arg);

typedef struct {
    orc_state *s;
    int nt;
    pthread_barrier_t start, done;
    task_fn fn;
    void *arg;
    int quit;
} pool_t;

typedef struct { pool_t *pool; int tid; } worker_arg;

static void *pool_worker(void *vp)
{
    worker_arg *w = (worker_arg *)vp;
    pool_t *pl = w->pool;
    for (;;) {
        pthread_barrier_wait(&pl->start);
        if (pl->quit) break;
        pl->fn(pl->s, w->tid, pl->nt, pl->arg);
        pthread_barrier_wait(&pl->done);
    }
    return NULL;
}

static pool_t *g_pool = NULL;

static void run_par(orc_state *s, task_fn fn, void *arg)
{
    if (s->nt <= 1 || !g_pool) {
        fn(s, 0, 1, arg);
        return;
    }
    g_pool->fn = fn;
    g_pool->arg = arg;
    pthread_barrier_wait(&g_pool->start);
    fn(s, 0, s->nt, arg);
    pthread_barrier_wait(&g_pool->done);
}

static void range_of(size_t n, int tid, int nt, size_t *b, size_t *e)
{
    *b = n * (size_t)tid / (size_t)nt;
    *e = n * (size_t)(tid + 1) / (size_t)nt;
}

/* ------------------------------------------------------------------ */
/* Step components                                                     */
/* ------------------------------------------------------------------ */

/* Lumped mass: solverExplicit.cpp:429-459 (GP-order accumulation) or per-element
 * sums + node gather (ORC_SUM_ELEMENT).  Stored per node. */
static void compute_mass(orc_state *s)
{
    size_t N = s->N, Ne = s->Ne, ng = s->ngp, nen = s->nen;
    for (size_t n = 0; n < N; n++) s->mass[n] = 0;
    if (s->p->sum_order == ORC_SUM_REFERENCE) {
        for (size_t e = 0; e < Ne; e++)
            for (size_t g = 0; g < ng; g++) {
                size_t gi = e * ng + g;
                double m[8];
                orc_gp_lumped_mass(s->nen, s->wJ[gi], s->rho[e], s->phi + gi * nen, m);
                for (size_t a = 0; a < nen; a++) s->mass[s->p->conn[e * nen + a]] += m[a];
            }
    } else {
        double *me = malloc(sizeof(double) * Ne * nen);
        for (size_t e = 0; e < Ne; e++) {
            for (size_t a = 0; a < nen; a++) me[e * nen + a] = 0;
            for (size_t g = 0; g < ng; g++) {
                size_t gi = e * ng + g;
                double m[8];
                orc_gp_lumped_mass(s->nen, s->wJ[gi], s->rho[e], s->phi + gi * nen, m);
                for (size_t a = 0; a < nen; a++) me[e * nen + a] += m[a];
            }
        }
        for (size_t n = 0; n < N; n++) {
            double acc = 0;
            for (int32_t k = s->rowptr[n]; k < s->rowptr[n + 1]; k++) acc += me[s->col[k]];
            s->mass[n] = acc;
        }
        free(me);
    }
}

/* Per-element work for ORC_SUM_ELEMENT: fe[e] = sum over GPs of the GP force. */
static void element_force(orc_state *s, size_t e, const double *u)
{
    size_t ng = s->ngp, nen = s->nen, N = s->N;
    const int32_t *cn = s->p->conn + e * nen;
    double uloc[24], F[9], P[9], f[24];
    for (size_t a = 0; a < nen; a++)
        for (int i = 0; i < 3; i++) uloc[a * 3 + i] = u[cn[a] + N * i];
    double *fe = s->fe + e * nen * 3;
    for (size_t k = 0; k < nen * 3; k++) fe[k] = 0;
    for (size_t g = 0; g < ng; g++) {
        size_t gi = e * ng + g;
        orc_kinematics(s->nen, s->dphi + gi * nen * 3, uloc, F);
        pk1_law(s->p->law, F, s->lam[e], s->mu[e], s->K1[e], P);
        orc_gp_force(s->nen, s->wJ[gi], P, s->dphi + gi * nen * 3, f);
        for (size_t a = 0; a < nen; a++)
            for (int i = 0; i < 3; i++) fe[a * 3 + i] += f[a + i * nen];
    }
}

static void task_elements(orc_state *s, int tid, int nt, void *arg)
{
    const double *u = (const double *)arg;
    size_t b, e;
    range_of(s->Ne, tid, nt, &b, &e);
    for (size_t k = b; k < e; k++) element_force(s, k, u);
}

static void task_gather(orc_state *s, int tid, int nt, void *arg)
{
    (void)arg;
    size_t b, e, N = s->N;
    range_of(N, tid, nt, &b, &e);
    for (size_t n = b; n < e; n++)
        for (int i = 0; i < 3; i++) {
            double acc = 0;
            for (int32_t k = s->rowptr[n]; k < s->rowptr[n + 1]; k++) acc += s->fe[(size_t)s->col[k] * 3 + i];
            s->fint[n + N * i] = acc;
        }
}

/* BCsGravity (boundaryConditions.cpp:1146-1172): force[n_a + 2N] += -w * rho * |J| * g * phi_a,
 * GP by GP in element order, onto what the loads before it have added. */
static void add_gravity(orc_state *s, const double *tg)
{
    size_t N = s->N, Ne = s->Ne, ng = s->ngp, nen = s->nen;
    for (size_t e = 0; e < Ne; e++)
        for (size_t g = 0; g < ng; g++) {
            const double t = tg[e * ng + g];
            const double *phi = s->phi + (e * ng + g) * nen;
            for (size_t a = 0; a < nen; a++) s->fext[s->p->conn[e * nen + a] + 2 * N] += t * phi[a];
        }
}

/* BCsPressureRamp / BCsPressureInst (boundaryConditions.cpp:370-496): per surface GP, F of the bulk element at
 * the point (classSpatialPoint::getF, classSpatialPoint.cpp:41-68: the node terms, then the identity),
 * n = F^-T n0 (inverse, transpose, multTensorVector), and force += w |J_s| phi_a J P n_i over the bulk
 * element's nodes and shape (ramp) or the face's (inst). */
static void add_pressure(orc_state *s, const orc_load *L, const struct load_state *l, const double *u, double t)
{
    const size_t N = s->N;
    const int nen = s->nen, ramp = (L->type == ORC_LOAD_PRESSURE_RAMP);
    double P = 0.;
    if (t >= L->t0 && t <= L->tf) P = ramp ? L->value / (L->tf - L->t0) * (t - L->t0) : L->value;
    for (int g = 0; g < l->ns; g++) {
        const int32_t *cn = s->p->conn + (size_t)l->se[g] * nen;
        const double *dphi = l->dphb + 24 * g;
        double F[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, Fi[9], FiT[9], n[3] = {0, 0, 0};
        for (int a = 0; a < nen; a++) {
            double disp[3];
            for (int i = 0; i < 3; i++) disp[i] = u ? u[cn[a] + N * i] : 0.0;
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) F[i * 3 + j] += dphi[a * 3 + j] * disp[i];
        }
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) F[i * 3 + j] += (i == j) ? 1 : 0;
        const double Jac = orc_det3(F);
        orc_inverse3(F, Fi);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) FiT[i * 3 + j] = Fi[j * 3 + i];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) n[i] += FiT[i * 3 + j] * l->n0[3 * g + j];
        const int na = ramp ? nen : (nen == 8 ? 4 : 3);
        for (int a = 0; a < na; a++) {
            const int32_t node = ramp ? cn[a] : l->sn[4 * g + a];
            const double ph = ramp ? l->phb[8 * g + a] : l->phs[4 * g + a];
            for (int i = 0; i < 3; i++) s->fext[node + N * i] += l->ws[g] * ph * Jac * P * n[i];
        }
    }
}

/* BCsTractionRamp / BCsTractionInst (boundaryConditions.cpp:218-271): per surface GP of the faces,
 * force += w |J_s| phi_a T_i over the face's nodes, T from getTraction (classNeumannBCs.h:48-99). */
static void add_traction(orc_state *s, const orc_load *L, const struct load_state *l, double t)
{
    const size_t N = s->N;
    const int nf = (s->nen == 8) ? 4 : 3, ramp = (L->type == ORC_LOAD_TRACTION_RAMP);
    double T[3] = {0, 0, 0};
    if (t >= L->t0 && t <= L->tf)
        for (int i = 0; i < 3; i++) T[i] = ramp ? L->force[i] / (L->tf - L->t0) * (t - L->t0) : L->force[i];
    for (int g = 0; g < l->ns; g++)
        for (int a = 0; a < nf; a++)
            for (int i = 0; i < 3; i++) s->fext[l->sn[4 * g + a] + N * i] += l->ws[g] * l->phs[4 * g + a] * T[i];
}

/* fext at time t for displacement u (NeumannBCManagement, boundaryConditions.cpp:75-140): zero, then every
 * block in input order -- nodal forces (BCsNodalForces, :1120-1134, with classNeumann::getTracAmount,
 * classNodes.h:101-108: F/(tf-t0) * t for t0 <= t <= tf, else 0), gravity and pressure. */
static void external_force(orc_state *s, const double *u, double t)
{
    const orc_problem *p = s->p;
    size_t N = s->N;
    for (size_t i = 0; i < 3 * N; i++) s->fext[i] = 0;
    for (int b = 0; b < p->n_loads; b++) {
        const orc_load *L = &p->loads[b];
        const struct load_state *l = &s->ld[b];
        if (L->type == ORC_LOAD_FORCE) {
            const int in = (t >= L->t0 && t <= L->tf);
            for (int k = 0; k < L->n; k++)
                for (int d = 0; d < 3; d++) s->fext[L->index[k] + N * d] += in ? l->fm[3 * k + d] * t : 0.0;
        } else if (L->type == ORC_LOAD_GRAVITY) {
            add_gravity(s, l->tg);
        } else if (L->type >= ORC_LOAD_TRACTION_RAMP) {
            add_traction(s, L, l, t);
        } else {
            add_pressure(s, L, l, u, t);
        }
    }
}

/* computeInternalForce: solverExplicit.cpp:363-379 (+ computeState, solvers.cpp:615-641) */
/* bulkViscosity (solverExplicit.cpp:513-542): Fdot = (F1 - F0) * (1/dt), trD = tr Fdot,
 * q = rho/J1 h (c1^2 h |trD| min(0, trD) + cl c0 trD), Cauchy = -q I, P = Cauchy F1^-T * (-J1). */
static void bulk_viscosity(const orc_problem *p, double h, double c0, const double *F1, const double *F0, double rho,
                           double dt, double *P)
{
    double Fdot[9];
    if (dt == 0) {
        for (int i = 0; i < 9; i++) Fdot[i] = 0.;
    } else {
        for (int i = 0; i < 9; i++) Fdot[i] = F1[i] + -1 * F0[i];
        for (int i = 0; i < 9; i++) Fdot[i] = Fdot[i] * (1 / dt);
    }
    double trD = 0;
    for (int i = 0; i < 3; i++) trD += Fdot[i * 3 + i];
    const double J1 = orc_det3(F1);
    const double Rho_1 = rho / J1;
    const double mn = (trD < 0.) ? trD : 0.; /* std::min(0., trD) */
    const double qBV = Rho_1 * h * (p->visc_c1 * p->visc_c1 * h * fabs(trD) * mn + p->visc_cl * c0 * trD);
    double C[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, Fi[9];
    for (int i = 0; i < 3; i++) C[i + i * 3] = -1 * qBV;
    orc_inverse3(F1, Fi);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double c = 0;
            for (int k = 0; k < 3; k++) c += C[i * 3 + k] * Fi[j * 3 + k];
            P[i * 3 + j] = c;
        }
    for (int i = 0; i < 9; i++) P[i] = P[i] * (-1 * J1);
}

/* computeInternalForce: solverExplicit.cpp:363-423 (+ computeState, solvers.cpp:615-641).  With bulk
 * viscosity, a second pass over the GPs adds wJ P_vis dphi with F0 the GP's F at the previous step, i.e.
 * at uprev.  The deviatoric variant (artificialViscosity, :469-503) never copies its stress into the PK1
 * (Cauchy_vis stays zero), so it adds nothing. */
static void internal_force_visc(orc_state *s, const double *u, const double *uprev, double dt)
{
    size_t N = s->N, Ne = s->Ne, ng = s->ngp, nen = s->nen;
    if (s->p->sum_order == ORC_SUM_REFERENCE) {
        for (size_t k = 0; k < 3 * N; k++) s->fint[k] = 0;
        double uloc[24], F[9], P[9], f[24];
        for (size_t e = 0; e < Ne; e++) {
            const int32_t *cn = s->p->conn + e * nen;
            for (size_t a = 0; a < nen; a++)
                for (int i = 0; i < 3; i++) uloc[a * 3 + i] = u[cn[a] + N * i];
            for (size_t g = 0; g < ng; g++) {
                size_t gi = e * ng + g;
                orc_kinematics(s->nen, s->dphi + gi * nen * 3, uloc, F);
                pk1_law(s->p->law, F, s->lam[e], s->mu[e], s->K1[e], P);
                orc_gp_force(s->nen, s->wJ[gi], P, s->dphi + gi * nen * 3, f);
                /* classGPs::getLocalDofs (classGPs.cpp:37-67): vDofs[a + n*i] = node_a + N*i */
                for (int i = 0; i < 3; i++)
                    for (size_t a = 0; a < nen; a++) s->fint[cn[a] + N * i] += f[a + i * nen];
            }
        }
        if (uprev && s->p->visc == ORC_VISC_BULK) {
            double uloc0[24], F0[9];
            for (size_t e = 0; e < Ne; e++) {
                const int32_t *cn = s->p->conn + e * nen;
                for (size_t a = 0; a < nen; a++)
                    for (int i = 0; i < 3; i++) {
                        uloc[a * 3 + i] = u[cn[a] + N * i];
                        uloc0[a * 3 + i] = uprev[cn[a] + N * i];
                    }
                const double h = s->p->visc_h > 0 ? s->p->visc_h : s->l0[e];
                const double c0 = s->p->visc_c0 > 0 ? s->p->visc_c0 : s->c[e];
                for (size_t g = 0; g < ng; g++) {
                    size_t gi = e * ng + g;
                    const double *dphi = s->dphi + gi * nen * 3;
                    orc_kinematics(s->nen, dphi, uloc, F);
                    orc_kinematics(s->nen, dphi, uloc0, F0);
                    bulk_viscosity(s->p, h, c0, F, F0, s->rho[e], dt, P);
                    for (size_t a = 0; a < nen; a++)
                        for (int i = 0; i < 3; i++)
                            for (int J = 0; J < 3; J++)
                                s->fint[cn[a] + N * i] += s->wJ[gi] * P[i * 3 + J] * dphi[a * 3 + J];
                }
            }
        }
    } else {
        run_par(s, task_elements, (void *)u);
        run_par(s, task_gather, NULL);
    }
}

static void internal_force(orc_state *s, const double *u) { internal_force_visc(s, u, NULL, 0.0); }

/* calculateCriticalTimeStep(nod, elem, uK): solvers.cpp:296-320 (u == NULL gives the
 * undeformed variant solvers.cpp:276-287, since min(l0, l0) = l0). */
static void task_dt(orc_state *s, int tid, int nt, void *arg)
{
    const double *u = (const double *)arg;
    size_t b, e;
    range_of(s->Ne, tid, nt, &b, &e);
    double best = 0;
    int have = 0;
    for (size_t k = b; k < e; k++) {
        double v;
        if (u) {
            double xyz[24];
            gather_xyz(s, (int)k, u, xyz);
            double dist = orc_char_length(s->nen, xyz);
            double mn = (dist < s->l0[k]) ? dist : s->l0[k]; /* std::min(minDist, dist) */
            v = mn / s->c[k];
        } else {
            v = s->l0[k] / s->c[k];
        }
        if (!have) { best = v; have = 1; }
        else if (v < best) best = v; /* std::min_element */
    }
    s->tmin[tid] = have ? best : INFINITY;
}

static double critical_dt(orc_state *s, const double *u)
{
    run_par(s, task_dt, (void *)u);
    double best = s->tmin[0];
    int nt = (s->nt > 1 && g_pool) ? s->nt : 1;
    for (int t = 1; t < nt; t++)
        if (s->tmin[t] < best) best = s->tmin[t];
    return best;
}

/* Energies: defoEnergy / kineticEnergy / volume with Sum operation,
 * dataExtraction.cpp:569-629, 631-707, 709-776 (GP order). */
static void task_energy(orc_state *s, int tid, int nt, void *arg)
{
    const double *const *uv = (const double *const *)arg;
    const double *u = uv[0], *v = uv[1];
    size_t b, e, N = s->N, ng = s->ngp, nen = s->nen;
    range_of(s->Ne, tid, nt, &b, &e);
    double ed = 0, ek = 0, vol = 0;
    double uloc[24], F[9];
    for (size_t k = b; k < e; k++) {
        const int32_t *cn = s->p->conn + k * nen;
        for (size_t a = 0; a < nen; a++)
            for (int i = 0; i < 3; i++) uloc[a * 3 + i] = u[cn[a] + N * i];
        for (size_t g = 0; g < ng; g++) {
            size_t gi = k * ng + g;
            orc_kinematics(s->nen, s->dphi + gi * nen * 3, uloc, F);
            double wJ = s->wJ[gi];
            double val = defo_energy_lm(s->p->law, F, s->lam[k], s->mu[k]);
            ed += val * wJ;
            double vel[3] = {0., 0., 0.};
            const double *phi = s->phi + gi * nen;
            for (size_t a = 0; a < nen; a++)
                for (int j = 0; j < 3; j++) vel[j] += phi[a] * v[cn[a] + j * N];
            double kv = 0;
            for (int j = 0; j < 3; j++) kv += 0.5 * s->rho[k] * vel[j] * vel[j];
            ek += kv * wJ;
            vol += wJ * orc_det3(F);
        }
    }
    s->tsum[tid * 3 + 0] = ed;
    s->tsum[tid * 3 + 1] = ek;
    s->tsum[tid * 3 + 2] = vol;
}

static void energies(orc_state *s, const double *u, const double *v, double *out)
{
    const double *uv[2] = {u, v};
    run_par(s, task_energy, (void *)uv);
    int nt = (s->nt > 1 && g_pool) ? s->nt : 1;
    out[0] = out[1] = out[2] = 0;
    for (int t = 0; t < nt; t++)
        for (int j = 0; j < 3; j++) out[j] += s->tsum[t * 3 + j];
}

/* ------------------------------------------------------------------ */
/* Public whole-solve entry points                                     */
/* ------------------------------------------------------------------ */

static int setup(orc_state *s, const orc_problem *p)
{
    memset(s, 0, sizeof(*s)); /* the callers free the state on every path, also after a rejected problem */
    if (!p || (p->nen != 8 && p->nen != 4) || p->n_nodes <= 0 || p->n_elem <= 0) return ORC_ERR_ARG;
    if (p->n_threads > 1 && p->sum_order != ORC_SUM_ELEMENT) return ORC_ERR_ARG;
    if (p->law < ORC_LAW_NH_COMPRESSIBLE || p->law > ORC_LAW_LINEAR) return ORC_ERR_ARG;
    if (p->visc < ORC_VISC_NONE || p->visc > ORC_VISC_DEV) return ORC_ERR_ARG;
    if (p->visc == ORC_VISC_BULK && p->sum_order != ORC_SUM_REFERENCE) return ORC_ERR_ARG;
    int rc = state_init(s, p);
    if (rc) return rc;
    s->nt = p->n_threads > 1 ? p->n_threads : 1;
    s->tmin = malloc(sizeof(double) * s->nt);
    s->tsum = malloc(sizeof(double) * 3 * s->nt);
    if (!s->tmin || !s->tsum) return ORC_ERR_MEM;
    return ORC_OK;
}

static pthread_t *pool_start(orc_state *s, worker_arg **wa)
{
    if (s->nt <= 1) return NULL;
    g_pool = calloc(1, sizeof(pool_t));
    g_pool->s = s;
    g_pool->nt = s->nt;
    pthread_barrier_init(&g_pool->start, NULL, s->nt);
    pthread_barrier_init(&g_pool->done, NULL, s->nt);
    pthread_t *th = malloc(sizeof(pthread_t) * s->nt);
    *wa = malloc(sizeof(worker_arg) * s->nt);
    for (int t = 1; t < s->nt; t++) {
        (*wa)[t].pool = g_pool;
        (*wa)[t].tid = t;
        pthread_create(&th[t], NULL, pool_worker, &(*wa)[t]);
    }
    return th;
}

static void pool_stop(orc_state *s, pthread_t *th, worker_arg *wa)
{
    if (!th) return;
    g_pool->quit = 1;
    pthread_barrier_wait(&g_pool->start);
    for (int t = 1; t < s->nt; t++) pthread_join(th[t], NULL);
    pthread_barrier_destroy(&g_pool->start);
    pthread_barrier_destroy(&g_pool->done);
    free(g_pool);
    g_pool = NULL;
    free(th);
    free(wa);
}

int32_t orc_lumped_mass(const orc_problem *p, double *mass)
{
    orc_state s;
    int rc = setup(&s, p);
    if (rc == ORC_OK) {
        compute_mass(&s);
        memcpy(mass, s.mass, sizeof(double) * (size_t)s.N);
    }
    state_free(&s);
    return rc;
}

int32_t orc_internal_force_visc(const orc_problem *p, const double *u, const double *uprev, double dt, double *fint)
{
    orc_state s;
    int rc = setup(&s, p);
    if (rc == ORC_OK) {
        internal_force_visc(&s, u, uprev, dt);
        memcpy(fint, s.fint, sizeof(double) * 3 * (size_t)s.N);
    }
    state_free(&s);
    return rc;
}

int32_t orc_internal_force(const orc_problem *p, const double *u, double *fint)
{
    orc_state s;
    int rc = setup(&s, p);
    worker_arg *wa = NULL;
    pthread_t *th = NULL;
    if (rc == ORC_OK) {
        th = pool_start(&s, &wa);
        internal_force(&s, u);
        memcpy(fint, s.fint, sizeof(double) * 3 * (size_t)s.N);
        pool_stop(&s, th, wa);
    }
    state_free(&s);
    return rc;
}

/* DEFO_ENERGY, KIN_ENERGY and VOLUME with the Sum operation for a given u and v (dataExtraction.cpp:569-776):
 * out[0] = sum_gp wJ W(F), out[1] = sum_gp wJ rho/2 |sum_a phi_a v_a|^2, out[2] = sum_gp wJ det F. */
int32_t orc_energies(const orc_problem *p, const double *u, const double *v, double *out)
{
    orc_state s;
    int rc = setup(&s, p);
    worker_arg *wa = NULL;
    pthread_t *th = NULL;
    if (rc == ORC_OK) {
        th = pool_start(&s, &wa);
        energies(&s, u, v, out);
        pool_stop(&s, th, wa);
    }
    state_free(&s);
    return rc;
}

/* State quantities at one GP as updateConstitutive leaves them (hyperElasticNeoHookean.cpp:70-126, 221-257;
 * hyperElasticStVenantKirchhoff.cpp:203-227; linearElastic.cpp:63-95) with computeCauchy, computeVMS,
 * computeVolumetricStrain, computeEquivalentStrain (maths.cpp:1241-1299).  q[0..8] Cauchy, q[9..17] GL strain
 * (small strain for the linear law), q[18] VMS, q[19] volumetric strain, q[20] equivalent strain; quantities a
 * law does not set keep the zero of the classStates constructor (classStates.cpp:15-44). */
static void gp_state_outputs(int32_t law, const double *F, double lam, double mu, double K1, double *q)
{
    double P[9], *cau = q, *E = q + 9;
    for (int i = 0; i < 21; i++) q[i] = 0;
    if (law == ORC_LAW_NH_BELYTSCHKO) {
        green_lagrange(F, E);
        pk1_belytschko_sig(F, mu, K1, P, cau);
    } else {
        if (law == ORC_LAW_SVK) green_lagrange(F, E);
        if (law == ORC_LAW_LINEAR)
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) E[i * 3 + j] = 0.5 * (F[i * 3 + j] + F[j * 3 + i] - 2. * (i == j ? 1.0 : 0.0));
        pk1_law(law, F, lam, mu, K1, P);
        double detJinv = 1. / orc_det3(F);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                double c = 0;
                for (int k = 0; k < 3; k++) c += P[i * 3 + k] * F[j * 3 + k];
                cau[i * 3 + j] = c;
            }
        for (int i = 0; i < 9; i++) cau[i] = cau[i] * detJinv;
    }
    q[18] = sqrt(0.5 * ((cau[0] - cau[4]) * (cau[0] - cau[4]) + (cau[4] - cau[8]) * (cau[4] - cau[8]) +
                        (cau[8] - cau[0]) * (cau[8] - cau[0]) + 6 * (cau[1] * cau[1] + cau[2] * cau[2] + cau[5] * cau[5])));
    if (law == ORC_LAW_NH_BELYTSCHKO || law == ORC_LAW_SVK) {
        double v = 0;
        for (int i = 0; i < 3; i++) v += E[i + i * 3];
        q[19] = v;
        q[20] = 0.6666 * sqrt(1.5 * (E[0] * E[0] + E[4] * E[4] + E[8] * E[8]) + 3 * (E[1] * E[1] + E[2] * E[2] + E[5] * E[5]));
    }
}

int32_t orc_external_force(const orc_problem *p, const double *u, double t, double *fext)
{
    orc_state s;
    int rc = setup(&s, p);
    if (rc == ORC_OK) {
        external_force(&s, u, t);
        memcpy(fext, s.fext, sizeof(double) * 3 * (size_t)s.N);
    }
    state_free(&s);
    return rc;
}

int32_t orc_pressure_face(const orc_problem *p, int32_t elem, int32_t face, double *xyz, double *ws, double *n0,
                          double *phi_s, double *phi_b, double *dphi_b)
{
    orc_state s;
    int32_t sn[16];
    int rc = setup(&s, p);
    if (rc == ORC_OK) rc = surface_gps(&s, elem, face, sn, xyz, ws, n0, phi_s, phi_b, dphi_b);
    state_free(&s);
    return rc;
}

int32_t orc_element_outputs(const orc_problem *p, const double *u, double *out)
{
    orc_state s;
    int rc = setup(&s, p);
    if (rc == ORC_OK) {
        size_t ng = s.ngp, nen = s.nen, N = s.N;
        for (size_t e = 0; e < (size_t)s.Ne; e++) {
            const int32_t *cn = p->conn + e * nen;
            double uloc[24], F[9], q[21], acc[21] = {0};
            for (size_t a = 0; a < nen; a++)
                for (int i = 0; i < 3; i++) uloc[a * 3 + i] = u[cn[a] + N * i];
            for (size_t g = 0; g < ng; g++) {  /* classElements::sumCauchy & co., GP order */
                orc_kinematics(s.nen, s.dphi + (e * ng + g) * nen * 3, uloc, F);
                gp_state_outputs(p->law, F, s.lam[e], s.mu[e], s.K1[e], q);
                for (int i = 0; i < 21; i++) acc[i] += q[i];
            }
            for (int i = 0; i < 21; i++) out[e * 21 + i] = acc[i] / (double)ng;  /* / cellCount (output.cpp:624-) */
        }
    }
    state_free(&s);
    return rc;
}

double orc_critical_dt(const orc_problem *p, const double *u)
{
    orc_state s;
    double dt = NAN;
    if (setup(&s, p) == ORC_OK) dt = p->scale * critical_dt(&s, u);
    state_free(&s);
    return dt;
}

/* The energy entries of history row k for the state u, v of the step just completed: DEFO_ENERGY, KIN_ENERGY,
 * VOLUME (dataExtraction.cpp:569-776), EXTERNAL_ENERGY = u . fext of the step (:867-875; s->fext holds it)
 * and the lumped-mass kinetic energy 1/2 sum m v^2 (sum over DOFs in index order). */
static void history_energies(orc_state *s, const double *u, const double *v, orc_history *hist, int64_t k)
{
    double en[3];
    energies(s, u, v, en);
    if (hist->e_def) hist->e_def[k] = en[0];
    if (hist->e_kin) hist->e_kin[k] = en[1];
    if (hist->volume) hist->volume[k] = en[2];
    size_t N = s->N;
    if (hist->e_ext) {
        double e = 0.0;
        for (size_t i = 0; i < 3 * N; i++) e += u[i] * s->fext[i];
        hist->e_ext[k] = e;
    }
    if (hist->e_kin_lumped) {
        double e = 0.0;
        for (size_t i = 0; i < 3 * N; i++) e += 0.5 * s->mass[i % N] * v[i] * v[i];
        hist->e_kin_lumped[k] = e;
    }
}

/* classSolverExplicit::calculate: solverExplicit.cpp:63-361 (sequential branch). */
int32_t orc_run(const orc_problem *p, double *u, double *v, double *a, orc_history *hist,
                orc_report *rep)
{
    orc_state s;
    int rc = setup(&s, p);
    orc_report r = {0, 0, 0, ORC_OK, 0};
    worker_arg *wa = NULL;
    pthread_t *th = NULL;
    double *uzero = NULL, *uprev = NULL, *incr = NULL, *fint_prev = NULL, *fext_prev = NULL;
    double w_ext = 0.0, w_react = 0.0;
    if (rc != ORC_OK) goto out;
    th = pool_start(&s, &wa);
    size_t N = s.N;

    compute_mass(&s); /* computeLumpedMassVector, :155 */

    /* Time specifications, :158-169 */
    double timeRun = p->t0;
    double totalTime = p->tf;
    double dt = critical_dt(&s, NULL) * p->scale;
    /* uK is fetched from the system only if timeRun < 1e-12 (:109, :176-178) */
    uzero = calloc(3 * N, sizeof(double));
    /* F of the previous state (solverExplicit.cpp:149: computeKinematicVariables(u0) + nextStep) */
    uprev = malloc(sizeof(double) * 3 * N);
    incr = calloc(3 * N, sizeof(double));
    fint_prev = calloc(3 * N, sizeof(double));
    fext_prev = calloc(3 * N, sizeof(double));
    if (!uzero || !uprev || !incr || !fint_prev || !fext_prev) { rc = ORC_ERR_MEM; goto out; }
    memcpy(uprev, u, sizeof(double) * 3 * N);
    const double *uK = (timeRun < 1E-12) ? u : uzero;
    /* a_prev is a at the last nextStep (:149 -> nonlinearSystem.cpp:710-724); it is
     * the same array as a until the solve overwrites a, so one array suffices. */
    int64_t nStep = 0;
    int every = p->energy_every;
    int64_t last_e = -1;
    for (;;) {
        if (p->n_steps > 0) {
            if (nStep >= p->n_steps) break;
            timeRun += dt;
        } else {
            if (!(timeRun < totalTime - 1E-12)) break;
            if (timeRun + dt > totalTime) { /* :227-232 */
                dt = totalTime - timeRun;
                timeRun = totalTime;
            } else {
                timeRun += dt;
            }
        }
        dt = critical_dt(&s, uK); /* :267 */
        dt = dt * p->scale;
        uK = u;
        if (dt < 1E-20) { rc = ORC_ERR_DT; break; } /* :270-273 */
        nStep++;
        /* dynamicPrediction, nonlinearSystem.cpp:1186-1201 (active DOFs) */
        for (size_t i = 0; i < 3 * N; i++) {
            incr[i] = 0.0;
            if (s.active[i]) {
                incr[i] = dt * v[i] + 0.5 * dt * dt * a[i];
                u[i] += incr[i];
            }
        }
        /* applyDisplacementBC, solverExplicit.cpp:545-575 (input order) */
        for (int k = 0; k < p->n_dir; k++) {
            int32_t ind = p->dir_dof[k];
            double U = 0;
            if (p->bc_mode == ORC_BC_INST) {
                U = (nStep == 1) ? p->dir_val[k] : 0.0; /* flagDiri resets after first use */
            } else if (p->bc_mode == ORC_BC_RAMP) {
                U = (p->dir_val[k] / (p->tf - p->t0)) * dt; /* UU = U/duration (classNodes.cpp:70) */
            } else { /* velocity amplitude (extension): V_n = V*min(n/ramp, 1) */
                double ramp = p->vel_ramp_steps > 0 ? (double)p->vel_ramp_steps : 1.0;
                double f = (double)nStep / ramp;
                if (f > 1.0) f = 1.0;
                double V = p->dir_val[k] * f;
                U = V * dt;
                incr[ind] = U;
                if (nStep >= 2) w_react += 0.5 * s.mass[(size_t)ind % N] * (V * V - v[ind] * v[ind]);
                u[ind] += U;
                v[ind] = V;
                a[ind] = 0.;
                continue;
            }
            incr[ind] = U;
            if (nStep >= 2) w_react += 0.5 * s.mass[(size_t)ind % N] * (U * U - v[ind] * v[ind]);
            u[ind] += U;
            v[ind] = U;
            a[ind] = 0.;
        }
        /* computeState + computeRightHandSide: fint(u) */
        internal_force_visc(&s, u, uprev, dt);
        /* NeumannBCManagement at timeRun (boundaryConditions.cpp:75-140) */
        external_force(&s, u, timeRun);
        /* work of the loads and of the supports over the step (trapezoid, from step 2 on) */
        if (nStep >= 2) {
            for (size_t i = 0; i < 3 * N; i++) {
                w_ext += 0.5 * (fext_prev[i] + s.fext[i]) * incr[i];
                if (!s.active[i])
                    w_react += 0.5 * ((fint_prev[i] - fext_prev[i]) + (s.fint[i] - s.fext[i])) * incr[i];
            }
        }
        memcpy(fint_prev, s.fint, sizeof(double) * 3 * N);
        memcpy(fext_prev, s.fext, sizeof(double) * 3 * N);
        /* systemSolve, nonlinearSystem.cpp:1096-1107: a = (-fint + fext)/m, v update */
        for (size_t i = 0; i < 3 * N; i++) {
            if (!s.active[i]) continue;
            double rhs = -s.fint[i] + s.fext[i];
            double m = s.mass[i % N];
            double anew = (m != 0.0) ? rhs / m : 0.0;
            v[i] += dt * ((1. - 0.5) * a[i] + 0.5 * anew);
            a[i] = anew;
        }
        if (p->visc == ORC_VISC_BULK) memcpy(uprev, u, sizeof(double) * 3 * N); /* nextStep: previous := current */
        if (!isfinite(u[1])) { rc = ORC_ERR_NAN; break; } /* :306-308 */
        int64_t k = nStep - 1;
        if (hist && k < hist->cap) {
            if (hist->time) hist->time[k] = timeRun;
            if (hist->dt) hist->dt[k] = dt;
            if (hist->w_ext) hist->w_ext[k] = w_ext;
            if (hist->w_react) hist->w_react[k] = w_react;
        }
        if (every > 0 && (nStep % every) == 0) {
            last_e = nStep;
            if (hist && k < hist->cap) history_energies(&s, u, v, hist, k);
        }
    }
    if (every > 0 && nStep > 0 && last_e != nStep && rc == ORC_OK && hist && nStep - 1 < hist->cap)
        history_energies(&s, u, v, hist, nStep - 1);
    r.steps = nStep;
    r.time = timeRun;
    r.dt_last = dt;
    r.status = rc;
out:
    pool_stop(&s, th, wa);
    free(uzero);
    free(uprev);
    free(incr);
    free(fint_prev);
    free(fext_prev);
    state_free(&s);
    if (rep) *rep = r;
    return rc;
}

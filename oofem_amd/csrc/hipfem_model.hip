// hfem_model: batched element/material evaluator for LSpace / LTRSpace /
// QSpace with IsoLE / MisesMat, standing in for the element loops of
// EngngModel::assemble (engngm.C:889-1037) and
// EngngModel::assembleVectorFromElements (engngm.C:1351-1535).
//
// Data layout in HBM
//   coords   nnode x 3 FP64 (node-major, gathered per element)
//   conn     nelem x nen int32 (element-major, the reference's local order)
//   eqn      nnode x 3 int32 (>0 free equation, <0 -prescribed equation)
//   eoff     nelem x nen x nen uint16: column offset of node b's first free
//            column inside every row of node a (all free rows of a node
//            share one column list), so entry (a,i | b,j) of K_e lands at
//            rowptr[eq(a,i)-1] + eoff + rank_b(j)
//   colors   element lists per color; no two elements of one color share a
//            node, so a color's launch scatters into K and f without atomics
//            and the result is deterministic (sum order = color order)
//   state    Gauss-point state, structure of arrays: field k of Gauss point g
//            at state[k * ngp_total + g], g = e * ngp + local point; fields
//            follow MisesMatStatus (misesmat.h:157-239): 0-5 committed plastic
//            strain, 6 committed kappa, 7-12 temp plastic strain, 13 temp
//            kappa, 14-19 trial deviatoric stress, 20-25 temp stress, 26-31
//            temp strain, 32-37 committed stress.
//
// Kernels
//   k_assemble<ET>   one 64-lane wave per element.  Lanes first compute the
//                    Gauss-point geometry (dN/dx, dV) into LDS, then each lane
//                    owns node pairs (a,b) and accumulates the 3x3 block
//                    B_a^T D B_b dV over the Gauss points in registers before
//                    scattering it.  IsoLE uses the closed form
//                    lambda n_i m_j + G (n_j m_i + delta_ij n.m); MisesMat
//                    the parts c I_dev + K 1x1 - m m^T of its tangent.
//   k_internal<ET>   one thread per Gauss point: geometry, strain B u_e, the
//                    material update (radial return for MisesMat), B^T sigma dV
//                    into LDS; then per element the Gauss-point sums in
//                    ascending order are stored, and k_gather_nodal sums them
//                    per node in ascending element order.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "hipfem_internal.h"
#include "hipfem_material.h"

namespace hipfem {

constexpr int NST = 38;  // state fields per Gauss point

// ------------------------------------------------------------ element rules
// Gauss rules: gaussintegrationrule.C:190-216, constants :1436-1463, tetra :492-502.
__device__ __forceinline__ void gauss_point(int et, int g, double xi[3], double &w)
{
    if (et == HIPFEM_LSPACE) {
        const double a = 0.577350269189626;
        xi[0] = (g & 4) ? a : -a;
        xi[1] = (g & 2) ? a : -a;
        xi[2] = (g & 1) ? a : -a;
        w = 1.0;
    } else if (et == HIPFEM_QSPACE) {
        const double x3[3] = {-0.774596669241483, 0.0, 0.774596669241483};
        const double w3[3] = {0.555555555555555, 0.888888888888888, 0.555555555555555};
        const int i = g / 9, j = (g / 3) % 3, k = g % 3;
        xi[0] = x3[i];
        xi[1] = x3[j];
        xi[2] = x3[k];
        w = w3[i] * w3[j] * w3[k];
    } else {
        xi[0] = xi[1] = xi[2] = 0.25;
        w = 1.0 / 6.0;
    }
}

// dN/dxi for node a: FEI3dHexaLin::evaldNdxi fei3dhexalin.C:127-175,
// FEI3dHexaQuad::evaldNdxi fei3dhexaquad.C:191-282.
__device__ __forceinline__ void dndxi_hexlin(int a, const double xi[3], double d[3])
{
    const double sx = (a == 2 || a == 3 || a == 6 || a == 7) ? 1.0 : -1.0;
    const double sy = (a == 1 || a == 2 || a == 5 || a == 6) ? 1.0 : -1.0;
    const double sz = (a < 4) ? 1.0 : -1.0;
    d[0] = sx * 0.125 * (1. + sy * xi[1]) * (1. + sz * xi[2]);
    d[1] = sy * 0.125 * (1. + sx * xi[0]) * (1. + sz * xi[2]);
    d[2] = sz * 0.125 * (1. + sx * xi[0]) * (1. + sy * xi[1]);
}

__device__ __forceinline__ void dndxi_hexquad(int a, const double xi[3], double d[3])
{
    const double u = xi[0], v = xi[1], w = xi[2];
    if (a < 8) {
        const double sx = (a == 2 || a == 3 || a == 6 || a == 7) ? 1.0 : -1.0;
        const double sy = (a == 1 || a == 2 || a == 5 || a == 6) ? 1.0 : -1.0;
        const double sz = (a < 4) ? 1.0 : -1.0;
        const double pu = 1. + sx * u, pv = 1. + sy * v, pw = 1. + sz * w;
        d[0] = 0.125 * sx * pv * pw * (2.0 * sx * u + sy * v + sz * w - 1.0);
        d[1] = 0.125 * sy * pu * pw * (2.0 * sy * v + sx * u + sz * w - 1.0);
        d[2] = 0.125 * sz * pu * pv * (2.0 * sz * w + sx * u + sy * v - 1.0);
    } else if (a < 16) {
        const double sw = a < 12 ? 1.0 : -1.0;
        const double pw = 1.0 + sw * w;
        switch (a & 3) {
        case 0:
            d[0] = -0.25 * (1.0 - v * v) * pw;
            d[1] = -0.5 * v * (1.0 - u) * pw;
            d[2] = sw * 0.25 * (1.0 - v * v) * (1.0 - u);
            break;
        case 1:
            d[0] = -0.5 * u * (1.0 + v) * pw;
            d[1] = 0.25 * (1.0 - u * u) * pw;
            d[2] = sw * 0.25 * (1.0 - u * u) * (1.0 + v);
            break;
        case 2:
            d[0] = 0.25 * (1.0 - v * v) * pw;
            d[1] = -0.5 * v * (1.0 + u) * pw;
            d[2] = sw * 0.25 * (1.0 - v * v) * (1.0 + u);
            break;
        default:
            d[0] = -0.5 * u * (1.0 - v) * pw;
            d[1] = -0.25 * (1.0 - u * u) * pw;
            d[2] = sw * 0.25 * (1.0 - u * u) * (1.0 - v);
            break;
        }
    } else {
        const int k = a - 16;
        const double cu = (k >= 2) ? 1.0 : -1.0, cv = (k == 1 || k == 2) ? 1.0 : -1.0;
        const double pu = 1.0 + cu * u, pv = 1.0 + cv * v;
        d[0] = cu * 0.25 * pv * (1.0 - w * w);
        d[1] = cv * 0.25 * pu * (1.0 - w * w);
        d[2] = -0.5 * pu * pv * w;
    }
}

template <int ET>
struct ElemT;
template <>
struct ElemT<HIPFEM_LSPACE> {
    static constexpr int NEN = 8, NGP = 8;
};
template <>
struct ElemT<HIPFEM_LTRSPACE> {
    static constexpr int NEN = 4, NGP = 1;
};
template <>
struct ElemT<HIPFEM_QSPACE> {
    static constexpr int NEN = 20, NGP = 27;
};

// Geometry at one Gauss point: dNdx (NEN x 3) and detJ.  Hexahedra:
// J = X dN/dxi, closed-form inverse (floatmatrix.C:768-815), dNdx = dN/dxi J^-1,
// detJ by FloatMatrix::giveDeterminant (floatmatrix.C:1085-1101).  Tetrahedra:
// cofactor form of FEI3dTetLin::evaldNdx (fei3dtetlin.C:116-168).
// X is NEN x 3 (row-major).  Returns false for a degenerate element.
template <int ET>
__device__ bool gp_geometry(const double *X, const double xi[3], double *dNdx, double &detJ)
{
    constexpr int NEN = ElemT<ET>::NEN;
    if constexpr (ET == HIPFEM_LTRSPACE) {
        const double x1 = X[0], y1 = X[1], z1 = X[2], x2 = X[3], y2 = X[4], z2 = X[5];
        const double x3 = X[6], y3 = X[7], z3 = X[8], x4 = X[9], y4 = X[10], z4 = X[11];
        detJ = ((x4 - x1) * (y2 - y1) * (z3 - z1) - (x4 - x1) * (y3 - y1) * (z2 - z1) +
                (x3 - x1) * (y4 - y1) * (z2 - z1) - (x2 - x1) * (y4 - y1) * (z3 - z1) +
                (x2 - x1) * (y3 - y1) * (z4 - z1) - (x3 - x1) * (y2 - y1) * (z4 - z1));
        if (!(detJ > 0.0)) return false;
        const double f = 1. / detJ;
        dNdx[0] = -((y3 - y2) * (z4 - z2) - (y4 - y2) * (z3 - z2)) * f;
        dNdx[3] = ((y4 - y3) * (z1 - z3) - (y1 - y3) * (z4 - z3)) * f;
        dNdx[6] = -((y1 - y4) * (z2 - z4) - (y2 - y4) * (z1 - z4)) * f;
        dNdx[9] = ((y2 - y1) * (z3 - z1) - (y3 - y1) * (z2 - z1)) * f;
        dNdx[1] = -((x4 - x2) * (z3 - z2) - (x3 - x2) * (z4 - z2)) * f;
        dNdx[4] = ((x1 - x3) * (z4 - z3) - (x4 - x3) * (z1 - z3)) * f;
        dNdx[7] = -((x2 - x4) * (z1 - z4) - (x1 - x4) * (z2 - z4)) * f;
        dNdx[10] = ((x3 - x1) * (z2 - z1) - (x2 - x1) * (z3 - z1)) * f;
        dNdx[2] = -((x3 - x2) * (y4 - y2) - (x4 - x2) * (y3 - y2)) * f;
        dNdx[5] = ((x4 - x3) * (y1 - y3) - (x1 - x3) * (y4 - y3)) * f;
        dNdx[8] = -((x1 - x4) * (y2 - y4) - (x2 - x4) * (y1 - y4)) * f;
        dNdx[11] = ((x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1)) * f;
        return true;
    } else {
        double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        double dN[NEN * 3];
#pragma unroll
        for (int a = 0; a < NEN; a++) {
            if constexpr (ET == HIPFEM_LSPACE)
                dndxi_hexlin(a, xi, dN + 3 * a);
            else
                dndxi_hexquad(a, xi, dN + 3 * a);
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) J[3 * i + j] += X[3 * a + i] * dN[3 * a + j];
        }
#define S(r, c) J[3 * ((r)-1) + ((c)-1)]
        const double det = S(1, 1) * S(2, 2) * S(3, 3) + S(1, 2) * S(2, 3) * S(3, 1) + S(1, 3) * S(2, 1) * S(3, 2) -
                           S(1, 3) * S(2, 2) * S(3, 1) - S(2, 3) * S(3, 2) * S(1, 1) - S(3, 3) * S(1, 2) * S(2, 1);
        if (!(fabs(det) > 1.e-30)) return false;
        double I[9];
        I[0] = (S(2, 2) * S(3, 3) - S(2, 3) * S(3, 2)) / det;
        I[3] = (S(2, 3) * S(3, 1) - S(2, 1) * S(3, 3)) / det;
        I[6] = (S(2, 1) * S(3, 2) - S(2, 2) * S(3, 1)) / det;
        I[1] = (S(1, 3) * S(3, 2) - S(1, 2) * S(3, 3)) / det;
        I[4] = (S(1, 1) * S(3, 3) - S(1, 3) * S(3, 1)) / det;
        I[7] = (S(1, 2) * S(3, 1) - S(1, 1) * S(3, 2)) / det;
        I[2] = (S(1, 2) * S(2, 3) - S(1, 3) * S(2, 2)) / det;
        I[5] = (S(1, 3) * S(2, 1) - S(1, 1) * S(2, 3)) / det;
        I[8] = (S(1, 1) * S(2, 2) - S(1, 2) * S(2, 1)) / det;
#undef S
        detJ = J[0] * J[4] * J[8] + J[3] * J[7] * J[2] + J[6] * J[1] * J[5] - J[2] * J[4] * J[6] -
               J[5] * J[7] * J[0] - J[8] * J[1] * J[3];
#pragma unroll
        for (int a = 0; a < NEN; a++)
#pragma unroll
            for (int j = 0; j < 3; j++)
                dNdx[3 * a + j] = dN[3 * a] * I[j] + dN[3 * a + 1] * I[3 + j] + dN[3 * a + 2] * I[6 + j];
        return true;
    }
}

// ------------------------------------------------------------ assembly kernel
template <int ET>
__global__ __launch_bounds__(64) void k_assemble(const int *__restrict__ elems, int nelem_color,
                                                 const double *__restrict__ coords, const int *__restrict__ conn,
                                                 const int *__restrict__ eqn, const int *__restrict__ emat,
                                                 const double *__restrict__ mat, const uint16_t *__restrict__ eoff,
                                                 const double *__restrict__ state, int ngp_total, int tangent,
                                                 const int32_t *__restrict__ rowptr, double *__restrict__ val,
                                                 int *__restrict__ bad, const int8_t *__restrict__ sflag,
                                                 const float *__restrict__ dsur, double *__restrict__ kbuf)
{
    // elems == nullptr: every element (block = element).  kbuf: the element's
    // 3x3 blocks are stored (9 per node pair (a, b), row-major) for
    // k_assemble_gather, else added into val (one color per launch)
    constexpr int NEN = ElemT<ET>::NEN, NGP = ElemT<ET>::NGP;
    __shared__ double sX[NEN * 3];
    __shared__ double sN[NGP][NEN * 3];
    __shared__ double sdV[NGP];
    __shared__ double sD[(ET == HIPFEM_LTRSPACE) ? 36 : NGP * 36];
    // MisesMat tangent parts D = c I_dev + K 1x1 - m m^T per point: c/2 and
    // K - c/3, m, and M g_a (M the symmetric tensor of m) per node
    __shared__ double sCK[NGP][2];
    __shared__ double sM[NGP][6];
    __shared__ double sV[NGP][NEN * 3];
    __shared__ int sEq[NEN * 3];
    __shared__ int sGeneral, sSur;

    const int ei = blockIdx.x;
    if (ei >= nelem_color) return;
    const int e = elems ? elems[ei] : ei;
    const int t = threadIdx.x;
    for (int k = t; k < NEN * 3; k += 64) {
        const int node = conn[(size_t)e * NEN + k / 3];
        sX[k] = coords[3 * (size_t)node + k % 3];
        sEq[k] = eqn[3 * (size_t)node + k % 3];
    }
    const double *mp = mat + (size_t)HIPFEM_MATP * emat[e];
    const int model = (int)mp[0];
    const double E = mp[1], nu = mp[2];
    if (t == 0) {
        sGeneral = (model == HIPFEM_MAT_MISES && tangent) ? 1 : 0;
        sSur = 0;   // set when a point of the element has a surrogate tangent
    }
    __syncthreads();
    for (int g = t; g < NGP; g += 64) {
        double xi[3], w, detJ;
        gauss_point(ET, g, xi, w);
        if (!gp_geometry<ET>(sX, xi, sN[g], detJ)) atomicMin(bad, e);
        sdV[g] = fabs(detJ) * w;
        if (sGeneral) {
            const size_t gp = (size_t)e * NGP + g;
            if (sflag && sflag[gp]) {
                // tangent predicted by the surrogate (upper triangle, row-major)
                double *D = sD + 36 * g;
                int k = 0;
                for (int r = 0; r < 6; r++)
                    for (int c = r; c < 6; c++) {
                        D[6 * r + c] = D[6 * c + r] = (double)dsur[(size_t)k * ngp_total + gp];
                        k++;
                    }
                sSur = 1;
            } else {
                double s[6];
#pragma unroll
                for (int i = 0; i < 6; i++) s[i] = state[(size_t)(14 + i) * ngp_total + gp];
                const double kn = state[(size_t)6 * ngp_total + gp], kt = state[(size_t)13 * ngp_total + gp];
                mises_tangent(E, nu, mp[3], mp[4], kn, kt, s, sD + 36 * g);
                double c, m[6];
                const double Kb = mises_tangent_parts(E, nu, mp[3], mp[4], kn, kt, s, c, m);
                sCK[g][0] = 0.5 * c;
                sCK[g][1] = Kb - c / 3.0;
#pragma unroll
                for (int i = 0; i < 6; i++) sM[g][i] = m[i];
            }
        }
    }
    __syncthreads();
    const bool parts = sGeneral && !sSur;   // every point's tangent in the parts form
    if (parts) {
        for (int k = t; k < NGP * NEN; k += 64) {
            const int g = k / NEN, a = k % NEN;
            const double n0 = sN[g][3 * a], n1 = sN[g][3 * a + 1], n2 = sN[g][3 * a + 2];
            const double *m = sM[g];   // M = [[m0 m5 m4] [m5 m1 m3] [m4 m3 m2]]
            sV[g][3 * a] = m[0] * n0 + m[5] * n1 + m[4] * n2;
            sV[g][3 * a + 1] = m[5] * n0 + m[1] * n1 + m[3] * n2;
            sV[g][3 * a + 2] = m[4] * n0 + m[3] * n1 + m[2] * n2;
        }
        __syncthreads();
    }
    const double G = E / (2.0 * (1. + nu));
    const double Kb = E / (3.0 * (1. - 2. * nu));
    const double lam = Kb - 2.0 * G / 3.0;
    // stored blocks (kbuf): the pairs a <= b, each also stored transposed as
    // (b, a), so every element matrix is exactly symmetric
    const int npairs = kbuf ? NEN * (NEN + 1) / 2 : NEN * NEN;
    for (int blk = t; blk < npairs; blk += 64) {
        int a = blk / NEN, b = blk % NEN;
        if (kbuf) {
            int rem = blk;
            a = 0;
            while (rem >= NEN - a) {
                rem -= NEN - a;
                a++;
            }
            b = a + rem;
        }
        double Kab[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (!sGeneral) {
            for (int g = 0; g < NGP; g++) {
                const double n0 = sN[g][3 * a], n1 = sN[g][3 * a + 1], n2 = sN[g][3 * a + 2];
                const double m0 = sN[g][3 * b], m1 = sN[g][3 * b + 1], m2 = sN[g][3 * b + 2];
                const double dv = sdV[g];
                const double nm = n0 * m0 + n1 * m1 + n2 * m2;
                const double n[3] = {n0, n1, n2}, m[3] = {m0, m1, m2};
#pragma unroll
                for (int i = 0; i < 3; i++)
#pragma unroll
                    for (int j = 0; j < 3; j++)
                        Kab[3 * i + j] += dv * (lam * n[i] * m[j] + G * (n[j] * m[i] + (i == j ? nm : 0.0)));
            }
        } else if (parts) {
            // B_a^T D B_b = c/2 ((g_a . g_b) I + g_b g_a^T) + (K - c/3) g_a g_b^T
            //               - (M g_a) (M g_b)^T
            for (int g = 0; g < NGP; g++) {
                const double n[3] = {sN[g][3 * a], sN[g][3 * a + 1], sN[g][3 * a + 2]};
                const double m[3] = {sN[g][3 * b], sN[g][3 * b + 1], sN[g][3 * b + 2]};
                const double va[3] = {sV[g][3 * a], sV[g][3 * a + 1], sV[g][3 * a + 2]};
                const double vb[3] = {sV[g][3 * b], sV[g][3 * b + 1], sV[g][3 * b + 2]};
                const double dv = sdV[g], ch = sCK[g][0], kl = sCK[g][1];
                const double nm = n[0] * m[0] + n[1] * m[1] + n[2] * m[2];
#pragma unroll
                for (int i = 0; i < 3; i++)
#pragma unroll
                    for (int j = 0; j < 3; j++)
                        Kab[3 * i + j] +=
                            dv * (ch * (m[i] * n[j] + (i == j ? nm : 0.0)) + kl * n[i] * m[j] - va[i] * vb[j]);
            }
        } else {
            for (int g = 0; g < NGP; g++) {
                const double *D = sD + 36 * g;
                const double n0 = sN[g][3 * a], n1 = sN[g][3 * a + 1], n2 = sN[g][3 * a + 2];
                const double m0 = sN[g][3 * b], m1 = sN[g][3 * b + 1], m2 = sN[g][3 * b + 2];
                const double dv = sdV[g];
                // columns of B_b: x (m0,0,0,0,m2,m1), y (0,m1,0,m2,0,m0), z (0,0,m2,m1,m0,0)
                double DB[6][3];
#pragma unroll
                for (int r = 0; r < 6; r++) {
                    DB[r][0] = D[6 * r + 0] * m0 + D[6 * r + 4] * m2 + D[6 * r + 5] * m1;
                    DB[r][1] = D[6 * r + 1] * m1 + D[6 * r + 3] * m2 + D[6 * r + 5] * m0;
                    DB[r][2] = D[6 * r + 2] * m2 + D[6 * r + 3] * m1 + D[6 * r + 4] * m0;
                }
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    Kab[0 + j] += dv * (n0 * DB[0][j] + n2 * DB[4][j] + n1 * DB[5][j]);
                    Kab[3 + j] += dv * (n1 * DB[1][j] + n2 * DB[3][j] + n0 * DB[5][j]);
                    Kab[6 + j] += dv * (n2 * DB[2][j] + n1 * DB[3][j] + n0 * DB[4][j]);
                }
            }
        }
        if (kbuf) {
            double *o = kbuf + ((size_t)e * NEN * NEN + a * NEN + b) * 9;
#pragma unroll
            for (int k = 0; k < 9; k++) o[k] = Kab[k];
            if (a != b) {
                double *ot = kbuf + ((size_t)e * NEN * NEN + b * NEN + a) * 9;
#pragma unroll
                for (int i = 0; i < 3; i++)
#pragma unroll
                    for (int j = 0; j < 3; j++) ot[3 * j + i] = Kab[3 * i + j];
            }
            continue;
        }
        const int off = eoff[((size_t)e * NEN + a) * NEN + b];
        if (off == 0xFFFF) continue;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const int ea = sEq[3 * a + i];
            if (ea <= 0) continue;
            const int base = rowptr[ea - 1] + off;
            int rnk = 0;
#pragma unroll
            for (int j = 0; j < 3; j++) {
                if (sEq[3 * b + j] <= 0) continue;
                val[base + rnk] += Kab[3 * i + j];
                rnk++;
            }
        }
    }
}

// The rows of node n from the stored element blocks.  The node's element
// list (element * NEN + local node a, ascending element) is taken 64 / NEN
// elements at a time; lane (slot, b) loads the block (a, b) of its element
// with its row offset (eoff) and column dofs, so that all loads of a pass are
// in flight together.  The slots then add into the LDS rows one after the
// other, in ascending element order, and each of the node's rows is written
// once.  One wave per node; the node's rows share one column list of length
// <= ASM_LMAX.
constexpr int ASM_LMAX = 256;
template <int NEN>
__global__ __launch_bounds__(64) void k_assemble_gather(int nnode, const int *__restrict__ nptr,
                                                        const int *__restrict__ nadj, const int *__restrict__ conn,
                                                        const int *__restrict__ eqn, const uint16_t *__restrict__ eoff,
                                                        const double *__restrict__ kbuf,
                                                        const int32_t *__restrict__ rowptr, double *__restrict__ val)
{
    static_assert(64 % NEN == 0, "element slots per pass");
    constexpr int SLOTS = 64 / NEN;
    __shared__ double sR[3][ASM_LMAX];
    const int n = blockIdx.x;
    if (n >= nnode) return;
    const int t = threadIdx.x, k = t / NEN, b = t % NEN;
    int q[3];
#pragma unroll
    for (int i = 0; i < 3; i++) q[i] = eqn[3 * (size_t)n + i];
    const int r0 = q[0] > 0 ? q[0] : q[1] > 0 ? q[1] : q[2];
    if (r0 <= 0) return;   // no free dof: no rows
    const int L = rowptr[r0] - rowptr[r0 - 1];
    for (int p = t; p < L; p += 64) sR[0][p] = sR[1][p] = sR[2][p] = 0.0;
    const int j0 = nptr[n], cnt = nptr[n + 1] - j0;
    for (int base = 0; base < cnt; base += SLOTS) {
        const int slot = base + k;
        double v[9];
        int pos[3] = {-1, -1, -1};
        if (slot < cnt) {
            const int ea = nadj[j0 + slot], e = ea / NEN, a = ea % NEN;
            const int off = eoff[((size_t)e * NEN + a) * NEN + b];
            const int nb = conn[(size_t)e * NEN + b];
            const double *src = kbuf + (((size_t)e * NEN + a) * NEN + b) * 9;
#pragma unroll
            for (int x = 0; x < 9; x++) v[x] = src[x];
            if (off != 0xFFFF) {
                int r = 0;
#pragma unroll
                for (int j = 0; j < 3; j++)
                    if (eqn[3 * (size_t)nb + j] > 0) pos[j] = off + r++;
            }
        }
        __syncthreads();
        for (int s = 0; s < SLOTS; s++) {
            if (k == s) {
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    if (q[i] <= 0) continue;
#pragma unroll
                    for (int j = 0; j < 3; j++)
                        if (pos[j] >= 0) sR[i][pos[j]] += v[3 * i + j];
                }
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
        if (q[i] <= 0) continue;
        double *row = val + rowptr[q[i] - 1];
        for (int p = t; p < L; p += 64) row[p] = sR[i][p];
    }
}

// ------------------------------------------------------ MisesMat surrogate
// Per-GP surrogate data of one internal-force evaluation.
struct SurArgs {
    const int *slot = nullptr;   // queue slot of the GP's prediction, -1 none (nullptr: surrogate off)
    const float *y = nullptr;    // predictions, 27 scaled values per slot
    float *dsur = nullptr;       // accepted tangent per GP, 21 x ngp_total (SoA, upper triangle)
    int8_t *flag = nullptr;      // 1: the GP's temp state and tangent come from the surrogate
    int8_t *rejected = nullptr;  // 1: rejected once in this load step (stays on the plain update)
    unsigned long long *counts = nullptr;   // [0] accepted, [1] rejected by the residual check
    float sig_scale = 1.f, tan_scale = 1.f;
    double tol = 0.0;
};

// The network's stress is accepted when the return-mapping equations of
// misesmat.C:181-263 hold at it, each within tol * sigma_y(kappa_n):
// the volumetric part is elastic (p = K tr(eps - epsp_n)), the deviator lies
// along the trial deviator, and the yield condition holds with the implied
// Delta kappa = (sigma_eq,trial - sigma_eq) / 3G > 0; and its tangent when
// every entry is within tol * E of the consistent tangent at that state
// (misesmat.C:496-544; an indefinite prediction would break the linear
// solver).  The temp state is then the one consistent with that stress:
// kappa_n + Delta kappa and epsp = eps - C^-1 sigma.
__device__ bool surrogate_update(double E, double nu, double sig0, double H, const double eps[6],
                                 const double epsp_n[6], double kappa_n, const float *y, float sig_scale,
                                 float tan_scale, double tol, MisesOut &o)
{
    const double G = E / (2.0 * (1. + nu));
    const double K = E / (3.0 * (1. - 2. * nu));
    double el[6];
#pragma unroll
    for (int i = 0; i < 6; i++) el[i] = eps[i] - epsp_n[i];
    const double mean = (el[0] + el[1] + el[2]) / 3.0;
    const double st[6] = {2. * G * (el[0] - mean), 2. * G * (el[1] - mean), 2. * G * (el[2] - mean),
                          G * el[3], G * el[4], G * el[5]};
    const double trialS = sqrt(st[0] * st[0] + st[1] * st[1] + st[2] * st[2] +
                               2. * (st[3] * st[3] + st[4] * st[4] + st[5] * st[5]));
    // network outputs are departures from the elastic trial
    double sp[6];
#pragma unroll
    for (int i = 0; i < 3; i++) sp[i] = st[i] + 3. * K * mean - (double)y[i] / (double)sig_scale;
#pragma unroll
    for (int i = 3; i < 6; i++) sp[i] = st[i] - (double)y[i] / (double)sig_scale;
    const double pp = (sp[0] + sp[1] + sp[2]) / 3.0;
    const double dv[6] = {sp[0] - pp, sp[1] - pp, sp[2] - pp, sp[3], sp[4], sp[5]};
    const double nrm = sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2] +
                            2. * (dv[3] * dv[3] + dv[4] * dv[4] + dv[5] * dv[5]));
    const double sy = sig0 + H * kappa_n;
    if (!(trialS > 0.0)) return false;
    const double dk = sqrt(1.5) * (trialS - nrm) / (3. * G);
    if (!(dk > 0.0)) return false;
    const double r1 = fabs(sqrt(1.5) * nrm - (sig0 + H * (kappa_n + dk)));
    const double f = nrm / trialS;
    double r2 = 0.0;
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const double d = dv[i] - f * st[i];
        r2 += (i < 3 ? 1.0 : 2.0) * d * d;
    }
    r2 = sqrt(r2);
    const double r3 = fabs(pp - K * 3.0 * mean);
    if (!(fmax(r1, fmax(r2, r3)) <= tol * sy)) return false;
    // the predicted tangent must match MisesMat's consistent tangent at the
    // accepted state (kappa_n + Delta kappa, trial deviator) within tol * E
    double D[36];
    mises_tangent(E, nu, sig0, H, kappa_n, kappa_n + dk, st, D);
    int k = 0;
    for (int r = 0; r < 6; r++)
        for (int c = r; c < 6; c++, k++)
            if (!(fabs(elastic_upper(E, nu, k) - (double)y[6 + k] / (double)tan_scale - D[6 * r + c]) <= tol * E))
                return false;
#pragma unroll
    for (int i = 0; i < 6; i++) {
        o.sig[i] = sp[i];
        o.sdev[i] = st[i];
    }
    const double ev = pp / (3. * K);
#pragma unroll
    for (int i = 0; i < 3; i++) o.epsp[i] = eps[i] - (dv[i] / (2. * G) + ev);
#pragma unroll
    for (int i = 3; i < 6; i++) o.epsp[i] = eps[i] - dv[i] / G;
    o.kappa = kappa_n + dk;
    return true;
}

// Queue the MisesMat Gauss points whose trial state lies more than band *
// sigma_y(kappa_n) outside the yield surface (the others stay with the
// plain update: elastic, or within the band around the surface) and write
// their scaled network inputs.  One thread per Gauss point, all elements.
template <int ET>
__global__ __launch_bounds__(256) void k_sur_classify(int nelem, const double *__restrict__ coords,
                                                      const int *__restrict__ conn, const int *__restrict__ eqn,
                                                      const int *__restrict__ emat, const double *__restrict__ mat,
                                                      const double *__restrict__ u_free,
                                                      const double *__restrict__ u_presc,
                                                      const double *__restrict__ state, int ngp_total, double band,
                                                      float in_scale, const int8_t *__restrict__ rejected,
                                                      int *__restrict__ slot, int *__restrict__ count,
                                                      float *__restrict__ X)
{
    constexpr int NEN = ElemT<ET>::NEN, NGP = ElemT<ET>::NGP, NC = 3 * NEN;
    const int gp = blockIdx.x * blockDim.x + threadIdx.x;
    if (gp >= ngp_total) return;
    const int e = gp / NGP, g = gp % NGP;
    const double *mp = mat + (size_t)HIPFEM_MATP * emat[e];
    if ((int)mp[0] != HIPFEM_MAT_MISES || rejected[gp]) {
        slot[gp] = -1;
        return;
    }
    double X3[NC], ue[NC];
    for (int a = 0; a < NEN; a++) {
        const int node = conn[(size_t)e * NEN + a];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            X3[3 * a + d] = coords[3 * (size_t)node + d];
            const int q = eqn[3 * (size_t)node + d];
            ue[3 * a + d] = (q > 0 && u_free) ? u_free[q - 1] : ((q < 0 && u_presc) ? u_presc[-q - 1] : 0.0);
        }
    }
    double xi[3], w, detJ, dNdx[NC];
    gauss_point(ET, g, xi, w);
    if (!gp_geometry<ET>(X3, xi, dNdx, detJ)) {
        slot[gp] = -1;
        return;
    }
    double eps[6] = {0, 0, 0, 0, 0, 0};
    for (int a = 0; a < NEN; a++) {
        const double nx = dNdx[3 * a], ny = dNdx[3 * a + 1], nz = dNdx[3 * a + 2];
        const double ux = ue[3 * a], uy = ue[3 * a + 1], uz = ue[3 * a + 2];
        eps[0] += nx * ux;
        eps[1] += ny * uy;
        eps[2] += nz * uz;
        eps[3] += nz * uy + ny * uz;
        eps[4] += nz * ux + nx * uz;
        eps[5] += ny * ux + nx * uy;
    }
    const double E = mp[1], nu = mp[2], sig0 = mp[3], H = mp[4];
    const double G = E / (2.0 * (1. + nu));
    double el[6], epsp_n[6];
#pragma unroll
    for (int i = 0; i < 6; i++) {
        epsp_n[i] = state[(size_t)i * ngp_total + gp];
        el[i] = eps[i] - epsp_n[i];
    }
    const double kappa_n = state[(size_t)6 * ngp_total + gp];
    const double mean = (el[0] + el[1] + el[2]) / 3.0;
    const double s0 = 2. * G * (el[0] - mean), s1 = 2. * G * (el[1] - mean), s2 = 2. * G * (el[2] - mean);
    const double s3 = G * el[3], s4 = G * el[4], s5 = G * el[5];
    const double trialS = sqrt(s0 * s0 + s1 * s1 + s2 * s2 + 2. * (s3 * s3 + s4 * s4 + s5 * s5));
    const double sy = sig0 + H * kappa_n;
    if (sqrt(1.5) * trialS - sy > band * sy) {
        const int k = atomicAdd(count, 1);
        slot[gp] = k;
        float *x = X + (size_t)k * 13;
#pragma unroll
        for (int i = 0; i < 6; i++) {
            x[i] = (float)(el[i] * in_scale);
            x[6 + i] = (float)(epsp_n[i] * in_scale);
        }
        x[12] = (float)(kappa_n * in_scale);
    } else {
        slot[gp] = -1;
    }
}

// -------------------------------------------------------- internal forces
template <int ET>
struct IntCfg {
    static constexpr int NGP = ElemT<ET>::NGP;
    static constexpr int EPB = (NGP >= 64) ? 1 : 64 / NGP;  // elements per 64-thread block
};

template <int ET>
__global__ __launch_bounds__(64) void k_internal(const int *__restrict__ elems, int nelem_color,
                                                 const double *__restrict__ coords, const int *__restrict__ conn,
                                                 const int *__restrict__ eqn, const int *__restrict__ emat,
                                                 const double *__restrict__ mat, const double *__restrict__ u_free,
                                                 const double *__restrict__ u_presc, double *__restrict__ state,
                                                 int ngp_total, int update_state, double *__restrict__ f_free,
                                                 double *__restrict__ f_presc, double *__restrict__ fbuf,
                                                 double *__restrict__ partials, int npart, int *__restrict__ bad,
                                                 int op, SurArgs sur)
{
    // elems == nullptr: every element, ei = e.  fbuf: the element force
    // vectors are stored (NC per element) for k_gather_nodal, else added to
    // f_free / f_presc (one color per launch)
    constexpr int NEN = ElemT<ET>::NEN, NGP = ElemT<ET>::NGP, EPB = IntCfg<ET>::EPB, NC = 3 * NEN;
    __shared__ double sF[EPB][NGP][NC];
    __shared__ double sh[3];
    __shared__ unsigned int scount[2];
    if (threadIdx.x < 2) scount[threadIdx.x] = 0;
    __syncthreads();
    const int t = threadIdx.x;
    const int el = t / NGP, g = t % NGP;
    const int ei = blockIdx.x * EPB + el;
    const bool active = (el < EPB) && (ei < nelem_color);
    if (active) {
        const int e = elems ? elems[ei] : ei;
        double X[NC], ue[NC];
        for (int a = 0; a < NEN; a++) {
            const int node = conn[(size_t)e * NEN + a];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                X[3 * a + d] = coords[3 * (size_t)node + d];
                const int q = eqn[3 * (size_t)node + d];
                ue[3 * a + d] = (q > 0 && u_free) ? u_free[q - 1] : ((q < 0 && u_presc) ? u_presc[-q - 1] : 0.0);
            }
        }
        double xi[3], w, detJ, dNdx[NC];
        gauss_point(ET, g, xi, w);
        if (!gp_geometry<ET>(X, xi, dNdx, detJ)) atomicMin(bad, e);
        const double dV = fabs(detJ) * w;
        // strain = B u_e (Voigt xx yy zz yz xz xy, engineering shear)
        double eps[6] = {0, 0, 0, 0, 0, 0};
        for (int a = 0; a < NEN; a++) {
            const double nx = dNdx[3 * a], ny = dNdx[3 * a + 1], nz = dNdx[3 * a + 2];
            const double ux = ue[3 * a], uy = ue[3 * a + 1], uz = ue[3 * a + 2];
            eps[0] += nx * ux;
            eps[1] += ny * uy;
            eps[2] += nz * uz;
            eps[3] += nz * uy + ny * uz;
            eps[4] += nz * ux + nx * uz;
            eps[5] += ny * ux + nx * uy;
        }
        const double *mp = mat + (size_t)HIPFEM_MATP * emat[e];
        const double E = mp[1], nu = mp[2];
        const size_t gp = (size_t)e * NGP + g;
        double sig[6];
        if (op != 0) {
            // tangent product: sigma = D eps with D the elastic (op 1) or the
            // TangentStiffness (op 2) matrix of the current temp state, so that
            // sum_g B^T D B u_e dV = K_e u_e
            double D[36], s[6];
            if ((int)mp[0] == HIPFEM_MAT_MISES && op == 2) {
#pragma unroll
                for (int i = 0; i < 6; i++) s[i] = state[(size_t)(14 + i) * ngp_total + gp];
                mises_tangent(E, nu, mp[3], mp[4], state[(size_t)6 * ngp_total + gp],
                              state[(size_t)13 * ngp_total + gp], s, D);
            } else {
#pragma unroll
                for (int i = 0; i < 6; i++) s[i] = 0.0;
                mises_tangent(E, nu, 0.0, 0.0, 0.0, 0.0, s, D);   // dKappa = 0: elastic D
            }
#pragma unroll
            for (int r = 0; r < 6; r++) {
                double v = 0.0;
#pragma unroll
                for (int c = 0; c < 6; c++) v += D[6 * r + c] * eps[c];
                sig[r] = v;
            }
        } else if ((int)mp[0] == HIPFEM_MAT_MISES) {
            double epsp_n[6];
#pragma unroll
            for (int i = 0; i < 6; i++) epsp_n[i] = state[(size_t)i * ngp_total + gp];
            const double kappa_n = state[(size_t)6 * ngp_total + gp];
            MisesOut o;
            bool used = false;
            if (sur.slot) {
                const int sl = sur.slot[gp];
                if (sl >= 0) {
                    used = surrogate_update(E, nu, mp[3], mp[4], eps, epsp_n, kappa_n, sur.y + (size_t)27 * sl,
                                            sur.sig_scale, sur.tan_scale, sur.tol, o);
                    atomicAdd(&scount[used ? 0 : 1], 1u);
                }
                if (update_state) {
                    sur.flag[gp] = used ? 1 : 0;
                    // a point rejected once keeps the plain update for the rest of the
                    // step, so accept/reject switches cannot repeat inside Newton
                    if (sl >= 0 && !used) sur.rejected[gp] = 1;
                    if (used) {
                        const float *yy = sur.y + (size_t)27 * sl + 6;
#pragma unroll
                        for (int k = 0; k < 21; k++)
                            sur.dsur[(size_t)k * ngp_total + gp] =
                                (float)(elastic_upper(E, nu, k) - (double)yy[k] / (double)sur.tan_scale);
                    }
                }
            }
            if (!used) mises_return(E, nu, mp[3], mp[4], eps, epsp_n, kappa_n, o);
#pragma unroll
            for (int i = 0; i < 6; i++) sig[i] = o.sig[i];
            if (update_state) {
#pragma unroll
                for (int i = 0; i < 6; i++) {
                    state[(size_t)(7 + i) * ngp_total + gp] = o.epsp[i];
                    state[(size_t)(14 + i) * ngp_total + gp] = o.sdev[i];
                }
                state[(size_t)13 * ngp_total + gp] = o.kappa;
            }
        } else {
            // IsoLE: sigma = (2G I_dev6 + K I6_I6) eps
            const double G = E / (2.0 * (1. + nu));
            const double K = E / (3.0 * (1. - 2. * nu));
            const double lam = K - 2.0 * G / 3.0;
            const double tr = eps[0] + eps[1] + eps[2];
            sig[0] = lam * tr + 2.0 * G * eps[0];
            sig[1] = lam * tr + 2.0 * G * eps[1];
            sig[2] = lam * tr + 2.0 * G * eps[2];
            sig[3] = G * eps[3];
            sig[4] = G * eps[4];
            sig[5] = G * eps[5];
        }
        if (update_state && op == 0) {
#pragma unroll
            for (int i = 0; i < 6; i++) {
                state[(size_t)(20 + i) * ngp_total + gp] = sig[i];
                state[(size_t)(26 + i) * ngp_total + gp] = eps[i];
            }
        }
        // B^T sigma dV
        for (int a = 0; a < NEN; a++) {
            const double nx = dNdx[3 * a], ny = dNdx[3 * a + 1], nz = dNdx[3 * a + 2];
            sF[el][g][3 * a + 0] = (nx * sig[0] + nz * sig[4] + ny * sig[5]) * dV;
            sF[el][g][3 * a + 1] = (ny * sig[1] + nz * sig[3] + nx * sig[5]) * dV;
            sF[el][g][3 * a + 2] = (nz * sig[2] + ny * sig[3] + nx * sig[4]) * dV;
        }
    }
    __syncthreads();
    if (sur.counts && t == 0 && (scount[0] | scount[1])) {
        atomicAdd(&sur.counts[0], (unsigned long long)scount[0]);
        atomicAdd(&sur.counts[1], (unsigned long long)scount[1]);
    }
    double n3[3] = {0, 0, 0};
    for (int idx = t; idx < EPB * NC; idx += 64) {
        const int l = idx / NC, c = idx % NC;
        const int ej = blockIdx.x * EPB + l;
        if (ej >= nelem_color) continue;
        const int e = elems ? elems[ej] : ej;
        double f = 0.0;
        for (int gg = 0; gg < NGP; gg++) f += sF[l][gg][c];
        if (fbuf) {
            fbuf[(size_t)e * NC + c] = f;
        } else {
            const int node = conn[(size_t)e * NEN + c / 3];
            const int q = eqn[3 * (size_t)node + c % 3];
            if (q > 0 && f_free) f_free[q - 1] += f;
            if (q < 0 && f_presc) f_presc[-q - 1] += f;
        }
        n3[c % 3] += f * f;
    }
    // block partial of the EBE norms (fixed order)
    if (partials) {
        for (int k = 0; k < 3; k++) {
            double v = n3[k];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            if (t == 0) sh[k] = v;
        }
        if (t == 0)
            for (int k = 0; k < 3; k++) partials[(size_t)k * npart + blockIdx.x] = sh[k];
    }
}

// f per dof = the sum of the stored element force vectors over the elements
// at the node, FP64, in the fixed order of the node's list (nptr / nadj:
// element * nen + local node, ascending element); one thread per node
__global__ void k_gather_nodal(int nnode, int nen, const int *__restrict__ nptr, const int *__restrict__ nadj,
                               const int *__restrict__ eqn, const double *__restrict__ fbuf,
                               double *__restrict__ f_free, double *__restrict__ f_presc)
{
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < nnode; n += gridDim.x * blockDim.x) {
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        const int j1 = nptr[n + 1];
        int j = nptr[n];
        // four list entries' loads in flight before their sums (same order)
        for (; j + 4 <= j1; j += 4) {
            double v[4][3];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int ea = nadj[j + k];
                const double *f = fbuf + (size_t)(ea / nen) * 3 * nen + 3 * (ea % nen);
                v[k][0] = f[0];
                v[k][1] = f[1];
                v[k][2] = f[2];
            }
#pragma unroll
            for (int k = 0; k < 4; k++) {
                s0 += v[k][0];
                s1 += v[k][1];
                s2 += v[k][2];
            }
        }
        for (; j < j1; j++) {
            const int ea = nadj[j];
            const double *f = fbuf + (size_t)(ea / nen) * 3 * nen + 3 * (ea % nen);
            s0 += f[0];
            s1 += f[1];
            s2 += f[2];
        }
        const double sv[3] = {s0, s1, s2};
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const int q = eqn[3 * (size_t)n + d];
            if (q > 0 && f_free) f_free[q - 1] = sv[d];
            if (q < 0 && f_presc) f_presc[-q - 1] = sv[d];
        }
    }
}

// ------------------------------------------------- element-by-element operator
// Per-GP tangent descriptor (FP32, six per point, point-major: desc[6 gp + k]):
// the TangentStiffness D of misesmat.C:496-544 written as
// D = c I_dev + K 1x1 - m m^T with c = 2G + factor2 and
// m = sqrt(-factor1) s_trial (factor1 <= 0), so that
// D eps = c dev(eps) + K tr(eps) 1 - m (m . eps); elastic points have c = 2G,
// m = 0.  Fields: c, m0, m1, m3, m4, m5.  The bulk modulus K comes from the
// material table and m2 = -(m0 + m1), since s_trial is deviatoric; c = -1
// marks a point whose general D (the surrogate's) sits in the snapshot.
__global__ void k_tangent_desc(int ngp_total, int ngp, const int *__restrict__ emat, const double *__restrict__ mat,
                               const double *__restrict__ state, int mode, float *__restrict__ desc,
                               const int8_t *__restrict__ sflag, const float *__restrict__ dsur,
                               float *__restrict__ dsnap)
{
    const int gp = blockIdx.x * blockDim.x + threadIdx.x;
    if (gp >= ngp_total) return;
    if (mode == 1 && sflag && sflag[gp]) {
        // general D: the surrogate's tangent of this assembly, kept for the
        // operator (the live one changes with every state update)
        desc[6 * (size_t)gp] = -1.0f;
#pragma unroll
        for (int k = 0; k < 21; k++) dsnap[(size_t)k * ngp_total + gp] = dsur[(size_t)k * ngp_total + gp];
        return;
    }
    const double *mp = mat + (size_t)HIPFEM_MATP * emat[gp / ngp];
    const double E = mp[1], nu = mp[2];
    const double G = E / (2.0 * (1. + nu));
    double c = 2.0 * G, m[6] = {0, 0, 0, 0, 0, 0};
    if (mode == 1 && (int)mp[0] == HIPFEM_MAT_MISES) {
        const double kn = state[(size_t)6 * ngp_total + gp], kt = state[(size_t)13 * ngp_total + gp];
        const double dk = kt - kn;
        if (dk > 0.0) {
            double s[6];
#pragma unroll
            for (int i = 0; i < 6; i++) s[i] = state[(size_t)(14 + i) * ngp_total + gp];
            const double trialS = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2] + 2. * s[3] * s[3] +
                                       2. * s[4] * s[4] + 2. * s[5] * s[5]);
            const double sigmaY = mp[3] + mp[4] * kn;
            const double factor = -2. * sqrt(6.) * G * G / trialS;
            const double f1 = factor * sigmaY / ((mp[4] + 3. * G) * trialS * trialS);
            c += factor * dk;
            const double sc = sqrt(-f1);
#pragma unroll
            for (int i = 0; i < 6; i++) m[i] = sc * s[i];
        }
    }
    float *d = desc + 6 * (size_t)gp;
    d[0] = (float)c;
    d[1] = (float)m[0];
    d[2] = (float)m[1];
    d[3] = (float)m[3];
    d[4] = (float)m[4];
    d[5] = (float)m[5];
}

__global__ void k_fill_d(int n, double *__restrict__ a, const PcgState *__restrict__ st)
{
    if (st && halted(st)) return;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) a[i] = 0.0;
}

// K_e x_e of LSpace elements: element arithmetic in FP32, the sum over
// elements in FP64 (neighbouring element forces of a smooth field nearly
// cancel at a node, by a factor of about h/L, so an FP32 sum would lose that
// many digits).
// 8 threads per element (consecutive lanes of one wave), 32 elements per
// block.  Lane g stands for the Gauss point and for the node whose local
// (xi, eta, zeta) signs are the bits (4, 2, 1) of g (LSpace nodes 5 1 6 2 8
// 4 7 3 for g = 0..7).  The trilinear field v = sum_a N_a v_a equals
// sum_S c_S prod_{j in S} xi_j over the subsets S of {xi, eta, zeta}, with
// 8 c_S = sum_a v_a prod_{j in S} s_aj: a Walsh-Hadamard transform of the 8
// nodal values, done across the 8 lanes by three xor butterflies (lane S
// ends with c_S).  The lanes share the coefficients of x (formed per element
// in FP64 and rounded once, by k_mf_erel) and of u (transformed here, in
// FP32, relative to the lane-0 node: a rigid translation adds nothing to the
// strain) through LDS; J and du/dxi at a point are then 3 FMAs per entry.  The forces take
// the transposed route: each point's T = dV sigma J^-T / 8 gives 8 subset
// terms F_S, whose sum over the points is a reduce-scatter butterfly (xor 4,
// 2, 1: 12 + 6 + 3 exchanges, a fixed order) leaving F_g on lane g, and the
// inverse transform (three more butterflies) gives lane g its node's force:
// either y += f over one color (elems = the color's list), or, with ebuf, the
// element's 24 forces are stored (elems = NULL: every element, one launch)
// for k_mf_gather to sum per node, or, with yslot, each force is stored in
// the element color's slot vector (yslot[color * neq + equation]) for
// k_mf_slot_sum.
constexpr int MF_EPB = 32;

// exchanges inside groups of 8 lanes with no address arithmetic: xor 1 and 2
// by DPP quad permutes, xor 4 and the group's first lane by ds_swizzle in
// bit-mask mode (source lane ((lane & and) | or) ^ xor within 32 lanes)
__device__ __forceinline__ float lane_xor(float v, int m)
{
    const int i = __float_as_int(v);
    if (m == 1) return __int_as_float(__builtin_amdgcn_mov_dpp(i, 0xB1, 0xF, 0xF, true));
    if (m == 2) return __int_as_float(__builtin_amdgcn_mov_dpp(i, 0x4E, 0xF, 0xF, true));
    return __int_as_float(__builtin_amdgcn_ds_swizzle(i, 0x1F | (4 << 10)));
}

__device__ __forceinline__ float lane_group_first(float v)
{
    return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x18));
}
// The element arithmetic of k_mf_hex for the 8 lanes of one element
// (lane g: Gauss point g and node nl(g)).  v[0..2]: the coefficient c_g of
// the element's coordinate field (zero on lane 0), v[3..5]: the operand
// values at the lane's node; sC: the
// element's 8x8 LDS coefficient slots; c / m / Kb: the lane point's
// tangent (c < 0: the general D of dgen at gp).  Every lane of the group
// takes part, and the caller's block-wide barrier orders the LDS slots.
// Returns in fr[0..2] the force of the lane's node.
__device__ __forceinline__ void mf_hex_forces(int g, float v[6], float (*sC)[8], float c, const float m[6],
                                              float Kb, const float *__restrict__ dgen, int ngp_total, size_t gp,
                                              float fr[24])
{
    // the 8 lanes of an element are consecutive lanes of one wave and every
    // lane takes part in the exchanges
#pragma unroll
    for (int d = 3; d < 6; d++) v[d] -= lane_group_first(v[d]);
#pragma unroll
    for (int mk = 4; mk >= 1; mk >>= 1) {
        const float sg = (g & mk) ? -1.f : 1.f;
#pragma unroll
        for (int k = 3; k < 6; k++) v[k] = fmaf(sg, lane_xor(v[k], mk), v[k]);
    }
#pragma unroll
    for (int k = 0; k < 3; k++) sC[g][k] = v[k];   // the coordinate coefficients come transformed
#pragma unroll
    for (int k = 3; k < 6; k++) sC[g][k] = 0.125f * v[k];
    __syncthreads();
    fr[0] = fr[1] = fr[2] = 0.f;   // F_0 = 0
    {
        const float a = 0.577350269189626f;
        const float xi = (g & 4) ? a : -a, eta = (g & 2) ? a : -a, ze = (g & 1) ? a : -a;
        const float ez = eta * ze, xz = xi * ze, xe = xi * eta;
        float C[8][6];
#pragma unroll
        for (int S = 1; S < 8; S++) {
            const float4 lo = *reinterpret_cast<const float4 *>(&sC[S][0]);
            const float2 hi = *reinterpret_cast<const float2 *>(&sC[S][4]);
            C[S][0] = lo.x; C[S][1] = lo.y; C[S][2] = lo.z; C[S][3] = lo.w; C[S][4] = hi.x; C[S][5] = hi.y;
        }
        float J[9], gu[9];   // J[3i+k] = dx_i/dxi_k, gu[3i+k] = du_i/dxi_k
#pragma unroll
        for (int i = 0; i < 3; i++) {
            J[3 * i + 0] = fmaf(C[7][i], ez, fmaf(C[5][i], ze, fmaf(C[6][i], eta, C[4][i])));
            J[3 * i + 1] = fmaf(C[7][i], xz, fmaf(C[3][i], ze, fmaf(C[6][i], xi, C[2][i])));
            J[3 * i + 2] = fmaf(C[7][i], xe, fmaf(C[3][i], eta, fmaf(C[5][i], xi, C[1][i])));
            gu[3 * i + 0] = fmaf(C[7][3 + i], ez, fmaf(C[5][3 + i], ze, fmaf(C[6][3 + i], eta, C[4][3 + i])));
            gu[3 * i + 1] = fmaf(C[7][3 + i], xz, fmaf(C[3][3 + i], ze, fmaf(C[6][3 + i], xi, C[2][3 + i])));
            gu[3 * i + 2] = fmaf(C[7][3 + i], xe, fmaf(C[3][3 + i], eta, fmaf(C[5][3 + i], xi, C[1][3 + i])));
        }
        const float det = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
                          J[2] * (J[3] * J[7] - J[4] * J[6]);
        const float id = __builtin_amdgcn_rcpf(det);   // v_rcp_f32 (1 ulp)
        float I[9];   // adj J = det J J^-1
        I[0] = J[4] * J[8] - J[5] * J[7];
        I[1] = J[2] * J[7] - J[1] * J[8];
        I[2] = J[1] * J[5] - J[2] * J[4];
        I[3] = J[5] * J[6] - J[3] * J[8];
        I[4] = J[0] * J[8] - J[2] * J[6];
        I[5] = J[2] * J[3] - J[0] * J[5];
        I[6] = J[3] * J[7] - J[4] * J[6];
        I[7] = J[1] * J[6] - J[0] * J[7];
        I[8] = J[0] * J[4] - J[1] * J[3];
        // H = du/dx det J = gu adj J; the strain takes the 1 / det J
        float H[9];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++)
                H[3 * i + j] = gu[3 * i] * I[j] + gu[3 * i + 1] * I[3 + j] + gu[3 * i + 2] * I[6 + j];
        const float eps[6] = {H[0] * id, H[4] * id, H[8] * id, (H[5] + H[7]) * id, (H[2] + H[6]) * id,
                              (H[1] + H[3]) * id};
        float sig[6];
        if (c >= 0.f) {
            const float tr = eps[0] + eps[1] + eps[2];
            const float me =
                m[0] * eps[0] + m[1] * eps[1] + m[2] * eps[2] + m[3] * eps[3] + m[4] * eps[4] + m[5] * eps[5];
            const float p = Kb * tr, t3 = tr * (1.f / 3.f);
            sig[0] = c * (eps[0] - t3) + p - m[0] * me;
            sig[1] = c * (eps[1] - t3) + p - m[1] * me;
            sig[2] = c * (eps[2] - t3) + p - m[2] * me;
            sig[3] = 0.5f * c * eps[3] - m[3] * me;
            sig[4] = 0.5f * c * eps[4] - m[4] * me;
            sig[5] = 0.5f * c * eps[5] - m[5] * me;
        } else {
            // general symmetric D (upper triangle, row-major) of a surrogate point
            float D[21];
#pragma unroll
            for (int k = 0; k < 21; k++) D[k] = dgen[(size_t)k * ngp_total + gp];
            sig[0] = D[0] * eps[0] + D[1] * eps[1] + D[2] * eps[2] + D[3] * eps[3] + D[4] * eps[4] + D[5] * eps[5];
            sig[1] = D[1] * eps[0] + D[6] * eps[1] + D[7] * eps[2] + D[8] * eps[3] + D[9] * eps[4] + D[10] * eps[5];
            sig[2] = D[2] * eps[0] + D[7] * eps[1] + D[11] * eps[2] + D[12] * eps[3] + D[13] * eps[4] + D[14] * eps[5];
            sig[3] = D[3] * eps[0] + D[8] * eps[1] + D[12] * eps[2] + D[15] * eps[3] + D[16] * eps[4] + D[17] * eps[5];
            sig[4] = D[4] * eps[0] + D[9] * eps[1] + D[13] * eps[2] + D[16] * eps[3] + D[18] * eps[4] + D[19] * eps[5];
            sig[5] = D[5] * eps[0] + D[10] * eps[1] + D[14] * eps[2] + D[17] * eps[3] + D[19] * eps[4] + D[20] * eps[5];
        }
        // T = dV sigma J^-T / 8 (T[3i+k]) = sign(det J) / 8 sigma adj J^T,
        // with dV = |det J| (the unit Gauss weight); f_a = sum_k T[., k] dN_a/dxi_k
        const float sc = det < 0.f ? -0.125f : 0.125f;
        const float s0 = sc * sig[0], s1 = sc * sig[1], s2 = sc * sig[2], s3 = sc * sig[3], s4 = sc * sig[4],
                    s5 = sc * sig[5];
        const float S[9] = {s0, s5, s4, s5, s1, s3, s4, s3, s2};
#pragma unroll
        for (int i = 0; i < 3; i++) {
            float T[3];
#pragma unroll
            for (int k = 0; k < 3; k++)
                T[k] = S[3 * i] * I[3 * k] + S[3 * i + 1] * I[3 * k + 1] + S[3 * i + 2] * I[3 * k + 2];
            // subset terms F_S (S as lane bits xi 4, eta 2, zeta 1; F_0 = 0)
            fr[3 + i] = T[2];
            fr[6 + i] = T[1];
            fr[9 + i] = fmaf(T[1], ze, T[2] * eta);
            fr[12 + i] = T[0];
            fr[15 + i] = fmaf(T[0], ze, T[2] * xi);
            fr[18 + i] = fmaf(T[0], eta, T[1] * xi);
            fr[21 + i] = fmaf(T[0], ez, fmaf(T[1], xz, T[2] * xe));
        }
    }
    // reduce-scatter over the 8 Gauss points: at the step with mask m the
    // thread keeps the half of its current components selected by (g & m)
    // and adds the partner's copy of that half; after m = 4, 2, 1 thread g
    // holds components 12 b2 + 6 b1 + 3 b0 + (0..2) = 3 g + (0..2), i.e. F_g
    {
        const bool h4 = g & 4, h2 = g & 2, h1 = g & 1;
#pragma unroll
        for (int j = 0; j < 12; j++) {
            const float keep = h4 ? fr[12 + j] : fr[j], send = h4 ? fr[j] : fr[12 + j];
            fr[j] = keep + lane_xor(send, 4);
        }
#pragma unroll
        for (int j = 0; j < 6; j++) {
            const float keep = h2 ? fr[6 + j] : fr[j], send = h2 ? fr[j] : fr[6 + j];
            fr[j] = keep + lane_xor(send, 2);
        }
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const float keep = h1 ? fr[3 + j] : fr[j], send = h1 ? fr[j] : fr[3 + j];
            fr[j] = keep + lane_xor(send, 1);
        }
        // inverse transform: f_a = sum_S prod_{j in S} s_aj F_S
#pragma unroll
        for (int mk = 4; mk >= 1; mk >>= 1) {
            const float sg = (g & mk) ? 1.f : -1.f;
#pragma unroll
            for (int i = 0; i < 3; i++) fr[i] = fmaf(sg, lane_xor(fr[i], mk), fr[i]);
        }
    }
}

// K_e x_e over the elements (or over elems[0..nel) of one color).  EL: the
// lanes' equations come from the per-lane words of eloc (first free
// equation and free-dof mask of the node), else from conn and eqn.  DESC:
// per-point tangent descriptors (c, m0, m1, m3, m4, m5; m2 = -(m0 + m1)),
// else the elastic D of the material table.  Every load is issued without a
// lane-dependent branch, those that need only the element first, so one
// wait covers them; x has at least one entry (the host launches nothing
// when there are no free equations).
template <bool EL, bool DESC>
__global__ __launch_bounds__(256, 8) void k_mf_hex(const int *__restrict__ elems, int nel,
                                                const float *__restrict__ erel, const int *__restrict__ conn,
                                                const int *__restrict__ eqn, const uint32_t *__restrict__ eloc,
                                                const int *__restrict__ emat,
                                                const double *__restrict__ mat, const float *__restrict__ desc,
                                                const float *__restrict__ dgen, int ngp_total,
                                                const float *__restrict__ x, double *__restrict__ y,
                                                float *__restrict__ ebuf, float *__restrict__ yslot,
                                                const uint8_t *__restrict__ ecolor, int neq,
                                                const PcgState *__restrict__ st)
{
    __shared__ __align__(16) float sC[MF_EPB][8][8];   // per element: c_S of x (0..2) and u (3..5)
    if (st && halted(st)) return;
    const int t = threadIdx.x, le = t >> 3, g = t & 7;
    // the lanes of a slot past the element count (last block only) repeat
    // the last element and store nothing (nel >= 1)
    const int ei0 = blockIdx.x * MF_EPB + le;
    const bool active = ei0 < nel;
    const int ei = active ? ei0 : nel - 1;
    const int nl = (0x26371504 >> (4 * g)) & 15;   // this lane's LSpace node (0-based)
    const int e = elems ? elems[ei] : ei;
    const size_t gp = (size_t)e * 8 + g;
    // loads that need only the element
    const int me = emat[e];
    float cx[3];   // lane g > 0: c_g of x (lane 0 reads c_1 and drops it)
#pragma unroll
    for (int d = 0; d < 3; d++) cx[d] = erel[(size_t)e * 21 + 3 * (g ? g - 1 : 0) + d];
    float2 d01, d23, d45;
    if constexpr (DESC) {
        d01 = *reinterpret_cast<const float2 *>(desc + 6 * gp);
        d23 = *reinterpret_cast<const float2 *>(desc + 6 * gp + 2);
        d45 = *reinterpret_cast<const float2 *>(desc + 6 * gp + 4);
    }
    int myq[3];   // equations of node nl (> 0 free)
    if constexpr (EL) {
        // one word per lane: the node's first free equation and its free
        // dofs, numbered consecutively (no dependent load of eqn)
        const uint32_t w = eloc[gp];
        const int base = (int)(w & 0x1FFFFFFFu) + 1;
        const uint32_t fm = w >> 29;
#pragma unroll
        for (int d = 0; d < 3; d++) myq[d] = (fm >> d & 1u) ? base + __builtin_popcount(fm & ((1u << d) - 1u)) : 0;
    } else {
        const int node = conn[(size_t)e * 8 + nl];
#pragma unroll
        for (int d = 0; d < 3; d++) myq[d] = eqn[3 * (size_t)node + d];
    }
    float v[6];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const float xv = x[myq[d] > 0 ? myq[d] - 1 : 0];
        v[d] = g ? cx[d] : 0.0f;
        v[3 + d] = myq[d] > 0 ? xv : 0.0f;
    }
    const double *mp = mat + (size_t)HIPFEM_MATP * me;
    const float E = (float)mp[1], nu = (float)mp[2];
    const float Kb = E * __builtin_amdgcn_rcpf(3.f * (1.f - 2.f * nu));
    float c, m[6];
    if constexpr (DESC) {
        c = d01.x;
        m[0] = d01.y;
        m[1] = d23.x;
        m[2] = -(d01.y + d23.x);
        m[3] = d23.y;
        m[4] = d45.x;
        m[5] = d45.y;
    } else {
        c = E * __builtin_amdgcn_rcpf(1.f + nu);
#pragma unroll
        for (int i = 0; i < 6; i++) m[i] = 0.f;
    }
    float fr[24];
    // every lane computes: the 8 lanes of an element are all active or all
    // inactive, and the exchanges stay inside the 8, so what inactive lanes
    // compute reaches no stored force
    mf_hex_forces(g, v, sC[le], c, m, Kb, dgen, ngp_total, gp, fr);
    if (active) {
        const size_t slot = yslot ? (size_t)ecolor[e] * neq : 0;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            if (ebuf)
                ebuf[(size_t)e * 24 + 3 * nl + i] = fr[i];
            else if (myq[i] > 0) {
                if (yslot)
                    yslot[slot + myq[i] - 1] = fr[i];
                else
                    y[myq[i] - 1] += (double)fr[i];
            }
        }
    }
}

// per-element geometry of k_mf_hex, built once per mesh: the coefficients
// c_S (S = 1..7) of the trilinear coordinate field, c_S = 1/8 sum_g x_g
// prod_{j in S} s_j(g) over the lanes g (node nl(g), local signs s_j(g) =
// +1 where bit j of g is set, bits (4, 2, 1) for xi, eta, zeta), formed in
// FP64 from coordinates relative to node 1 and rounded to FP32 once
__global__ void k_mf_erel(int nel, const double *__restrict__ coords, const int *__restrict__ conn,
                          float *__restrict__ erel)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nel) return;
    const int n0 = conn[(size_t)e * 8];
    double x[8][3];
    for (int g = 0; g < 8; g++) {
        const int n = conn[(size_t)e * 8 + ((0x26371504 >> (4 * g)) & 15)];
        for (int d = 0; d < 3; d++) x[g][d] = coords[3 * (size_t)n + d] - coords[3 * (size_t)n0 + d];
    }
    for (int S = 1; S < 8; S++)
        for (int d = 0; d < 3; d++) {
            double c = 0.0;
            for (int g = 0; g < 8; g++) c += (__builtin_popcount(S & ~g) & 1) ? -x[g][d] : x[g][d];
            erel[(size_t)e * 21 + 3 * (S - 1) + d] = (float)(0.125 * c);
        }
}

// y[i] = sum over the colors touching equation i, ascending, of the FP32
// slot values, in FP64
__global__ void k_mf_slot_sum(int neq, int ncolors, const uint16_t *__restrict__ eqmask,
                              const float *__restrict__ yslot, double *__restrict__ y, const PcgState *__restrict__ st)
{
    if (st && halted(st)) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < neq; i += gridDim.x * blockDim.x) {
        const unsigned m = eqmask[i];
        double s = 0.0;
        for (int c = 0; c < ncolors; c++)
            if (m >> c & 1u) s += (double)yslot[(size_t)c * neq + i];
        y[i] = s;
    }
}

// y = sum of the stored element forces per free equation, FP64, in the
// fixed order of the node's element list (nptr / nadj: element * 8 + local
// node, ascending element); one thread per node writes all its free dofs.
__global__ void k_mf_gather(int nnode, const int *__restrict__ nptr, const int *__restrict__ nadj,
                            const int *__restrict__ eqn, const float *__restrict__ ebuf, double *__restrict__ y,
                            const PcgState *__restrict__ st)
{
    if (st && halted(st)) return;
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < nnode; n += gridDim.x * blockDim.x) {
        const int q0 = eqn[3 * (size_t)n], q1 = eqn[3 * (size_t)n + 1], q2 = eqn[3 * (size_t)n + 2];
        if (q0 <= 0 && q1 <= 0 && q2 <= 0) continue;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        for (int j = nptr[n]; j < nptr[n + 1]; j++) {
            const int ea = nadj[j];
            const float *f = ebuf + (size_t)(ea >> 3) * 24 + 3 * (ea & 7);
            s0 += (double)f[0];
            s1 += (double)f[1];
            s2 += (double)f[2];
        }
        if (q0 > 0) y[q0 - 1] = s0;
        if (q1 > 0) y[q1 - 1] = s1;
        if (q2 > 0) y[q2 - 1] = s2;
    }
}

__global__ void k_fill(int64_t n, double v, double *__restrict__ a)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) a[i] = v;
}

__global__ void k_zero_f32(int64_t n, float *__restrict__ a)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) a[i] = 0.f;
}

__global__ __launch_bounds__(1024) void k_reduce3(const double *__restrict__ partials, int np, double *out)
{
    __shared__ double sh[16];
    for (int k = 0; k < 3; k++) {
        double s = 0.0;
        for (int i = threadIdx.x; i < np; i += 1024) s += partials[(size_t)k * np + i];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        if (threadIdx.x % 64 == 0) sh[threadIdx.x / 64] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            double tt = 0.0;
            for (int i = 0; i < 16; i++) tt += sh[i];
            out[k] = tt;
        }
        __syncthreads();
    }
}

// rhs = RT - F and the squared sums of rhs and RT per dof id (u, v, w), the
// quantities NRSolver::checkConvergence groups by dof id (nrsolver.C:648-660)
__global__ __launch_bounds__(256) void k_residual_groups(int neq, const int8_t *__restrict__ dofid,
                                                         const double *__restrict__ RT, const double *__restrict__ F,
                                                         double *__restrict__ rhs, double *__restrict__ partials,
                                                         int np)
{
    __shared__ double sh[6][4];
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (int i = blockIdx.x * 256 + threadIdx.x; i < neq; i += gridDim.x * 256) {
        const double rt = RT ? RT[i] : 0.0;
        const double r = rt - F[i];
        rhs[i] = r;
        const int d = dofid[i];
        s[d] += r * r;
        s[3 + d] += rt * rt;
    }
    for (int k = 0; k < 6; k++) {
        double v = s[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (threadIdx.x % 64 == 0) sh[k][threadIdx.x / 64] = v;
    }
    __syncthreads();
    if (threadIdx.x < 6) partials[(size_t)threadIdx.x * np + blockIdx.x] =
        sh[threadIdx.x][0] + sh[threadIdx.x][1] + sh[threadIdx.x][2] + sh[threadIdx.x][3];
}

// squared sums of x and y per dof id (u, v, w) over the free equations: the
// displacement-criterion sums of NRSolver::checkConvergence (nrsolver.C:648-660)
__global__ __launch_bounds__(256) void k_group_sums(int neq, const int8_t *__restrict__ dofid,
                                                    const double *__restrict__ x, const double *__restrict__ y,
                                                    double *__restrict__ partials, int np)
{
    __shared__ double sh[6][4];
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (int i = blockIdx.x * 256 + threadIdx.x; i < neq; i += gridDim.x * 256) {
        const int d = dofid[i];
        const double xi = x[i], yi = y[i];
        s[d] += xi * xi;
        s[3 + d] += yi * yi;
    }
    for (int k = 0; k < 6; k++) {
        double v = s[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (threadIdx.x % 64 == 0) sh[k][threadIdx.x / 64] = v;
    }
    __syncthreads();
    if (threadIdx.x < 6) partials[(size_t)threadIdx.x * np + blockIdx.x] =
        sh[threadIdx.x][0] + sh[threadIdx.x][1] + sh[threadIdx.x][2] + sh[threadIdx.x][3];
}

__global__ __launch_bounds__(1024) void k_reduce6(const double *__restrict__ partials, int np, double *out)
{
    __shared__ double sh[16];
    for (int k = 0; k < 6; k++) {
        double s = 0.0;
        for (int i = threadIdx.x; i < np; i += 1024) s += partials[(size_t)k * np + i];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        if (threadIdx.x % 64 == 0) sh[threadIdx.x / 64] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int i = 0; i < 16; i++) t += sh[i];
            out[k] = t;
        }
        __syncthreads();
    }
}

// MisesMatStatus::updateYourself misesmat.C:720-728 (and the stress of
// StructuralMaterialStatus::updateYourself)
__global__ void k_commit(int ngp_total, double *__restrict__ state)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngp_total) return;
    const size_t n = ngp_total;
    for (int i = 0; i < 6; i++) {
        state[i * n + g] = state[(7 + i) * n + g];
        state[(32 + i) * n + g] = state[(20 + i) * n + g];
    }
    state[6 * n + g] = state[13 * n + g];
}

// MisesMatStatus::initTempStatus (misesmat.C:707-715 with
// StructuralMaterialStatus::initTempStatus, structuralms.C:121-141): temp
// plastic strain, kappa and stress back to the committed values, trial
// deviator cleared.  The temp strain has no committed copy here; the next
// stress evaluation rewrites it before anything reads it.
__global__ void k_restart(int ngp_total, double *__restrict__ state)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngp_total) return;
    const size_t n = ngp_total;
    for (int i = 0; i < 6; i++) {
        state[(7 + i) * n + g] = state[i * n + g];
        state[(14 + i) * n + g] = 0.0;
        state[(20 + i) * n + g] = state[(32 + i) * n + g];
    }
    state[13 * n + g] = state[6 * n + g];
}

}  // namespace hipfem

using namespace hipfem;

struct hfem_model {
    hipfem_ctx *ctx = nullptr;
    int nnode = 0, etype = 0, nen = 0, ngp = 0, nelem = 0, nmat = 0, neq = 0, npeq = 0;
    std::vector<int> h_conn, h_eqn;
    double *d_coords = nullptr, *d_mat = nullptr, *d_state = nullptr;
    int *d_conn = nullptr, *d_emat = nullptr, *d_eqn = nullptr, *d_bad = nullptr;
    uint16_t *d_eoff = nullptr;
    long pattern_id = -1;
    int ncolors = 0;
    std::vector<int> color_ptr;
    int *d_colors = nullptr;
    int ngp_total = 0;
    // host-API staging
    double *d_uf = nullptr, *d_up = nullptr, *d_ff = nullptr, *d_fp = nullptr, *d_out3 = nullptr;
    int8_t *d_dofid = nullptr;   // dof id (0 u, 1 v, 2 w) of every free equation
    // element-by-element operator: per-GP FP32 tangent descriptors of the last
    // assembly and the matrices that assembly linked
    float *d_tdesc = nullptr;
    int *d_nptr = nullptr, *d_nadj = nullptr;   // node -> (element * 8 + local node), ascending element
    float *d_ebuf = nullptr;                     // element forces of the gather form (24 per element)
    bool tdesc_used = false;     // false: every point elastic, D from the material table
    bool has_mises = false;
    // element operator form: 0 one launch per color adding into the FP64
    // result, 1 gather (element forces stored, summed per node), 2 slots
    // (every element in one launch, each color's forces stored in its own
    // FP32 slot vector, summed per equation in color order)
    int mf_form = 2;
    std::vector<uint8_t> h_ecolor;               // color of every element
    std::vector<uint64_t> h_nodemask;            // colors of the elements at every node
    uint8_t *d_ecolor = nullptr;
    uint16_t *d_eqmask = nullptr;                // colors touching every free equation (slot form)
    float *d_yslot = nullptr;                    // ncolors x neq FP32 slots (slot form)
    float *d_erel = nullptr;                     // per element: coordinate coefficients c_1..c_7 (21 FP32)
    // internal forces: node -> (element * nen + local node), ascending
    // element, and the element force vectors of one evaluation
    int *d_fnptr = nullptr, *d_fnadj = nullptr;
    double *d_fbuf = nullptr;
    // stiffness assembly by rows (LSpace): the element blocks of one
    // assembly (64 x 9 doubles per element) and the longest row
    double *d_kbuf = nullptr;
    bool kbuf_failed = false;
    int max_rowlen = 0;
    // per element and k_mf_hex lane g: first free equation (0-based, bits
    // 0..28) and free-dof mask (bits 29..31) of the lane's node; built once
    // per mesh when every node numbers its free dofs consecutively
    uint32_t *d_eloc = nullptr;
    bool eloc_checked = false;
    std::vector<hipcsr_mtx *> linked;
    // MisesMat surrogate (hfem_model_set_surrogate)
    hfem_mlp *sur = nullptr;
    double sur_band = 0.05, sur_tol = 1e-4;
    float sur_scales[3] = {1, 1, 1};
    int *d_slot = nullptr, *d_qcount = nullptr;
    float *d_sx = nullptr, *d_sy = nullptr, *d_dsur = nullptr;
    float *d_dsnap = nullptr;    // surrogate tangents of the last assembly (what K holds)
    int8_t *d_sflag = nullptr, *d_srej = nullptr;
    unsigned long long *d_scounts = nullptr;   // accepted, rejected
    unsigned long long queued_total = 0;
};

static int model_free(hfem_model *md)
{
    for (hipcsr_mtx *k : md->linked) k->mf_model = nullptr;
    md->linked.clear();
    {
        void *sp[] = {md->d_slot, md->d_qcount, md->d_sx, md->d_sy, md->d_dsur, md->d_dsnap, md->d_sflag, md->d_srej,
                      md->d_scounts};
        for (void *p : sp)
            if (p) HIPFEM_CHECK(hipFree(p));
        md->d_slot = md->d_qcount = nullptr;
        md->d_sx = md->d_sy = md->d_dsur = md->d_dsnap = nullptr;
        md->d_sflag = md->d_srej = nullptr;
        md->d_scounts = nullptr;
        md->sur = nullptr;
    }
    void *ps[] = {md->d_coords, md->d_mat, md->d_state, md->d_conn, md->d_emat, md->d_eqn, md->d_bad,
                  md->d_eoff,   md->d_colors, md->d_uf, md->d_up, md->d_ff, md->d_fp, md->d_out3, md->d_dofid,
                  md->d_tdesc, md->d_nptr, md->d_nadj, md->d_ebuf, md->d_ecolor, md->d_eqmask, md->d_yslot,
                  md->d_erel, md->d_eloc, md->d_fnptr, md->d_fnadj, md->d_fbuf, md->d_kbuf};
    for (void *p : ps)
        if (p) HIPFEM_CHECK(hipFree(p));
    md->d_tdesc = nullptr;
    md->d_nptr = md->d_nadj = nullptr;
    md->d_ebuf = nullptr;
    md->d_ecolor = nullptr;
    md->d_eqmask = nullptr;
    md->d_yslot = nullptr;
    md->d_erel = nullptr;
    md->d_eloc = nullptr;
    md->eloc_checked = false;
    md->d_fnptr = md->d_fnadj = nullptr;
    md->d_fbuf = nullptr;
    md->d_kbuf = nullptr;
    md->kbuf_failed = false;
    md->max_rowlen = 0;
    md->d_coords = md->d_mat = md->d_state = md->d_uf = md->d_up = md->d_ff = md->d_fp = md->d_out3 = nullptr;
    md->d_conn = md->d_emat = md->d_eqn = md->d_bad = md->d_colors = nullptr;
    md->d_dofid = nullptr;
    md->d_eoff = nullptr;
    md->pattern_id = -1;
    md->ncolors = 0;
    return 0;
}

static int nen_of(int et) { return et == HIPFEM_LSPACE ? 8 : et == HIPFEM_LTRSPACE ? 4 : et == HIPFEM_QSPACE ? 20 : 0; }
static int ngp_of(int et) { return et == HIPFEM_LSPACE ? 8 : et == HIPFEM_LTRSPACE ? 1 : et == HIPFEM_QSPACE ? 27 : 0; }

extern "C" int hfem_model_create(hipfem_ctx *ctx, hfem_model **out)
{
    HIPFEM_REQUIRE(ctx && out, HIPFEM_EINVAL, "hfem_model_create: null argument");
    hfem_model *m = new hfem_model();
    m->ctx = ctx;
    ctx_retain(ctx);
    *out = m;
    return 0;
}

extern "C" void hfem_model_destroy(hfem_model *md)
{
    if (!md) return;
    hipfem_ctx *c = md->ctx;
    {
        DeviceGuard g(c->device);
        (void)hipStreamSynchronize(c->stream);
        model_free(md);
        delete md;
    }
    ctx_release(c);
}

extern "C" int hfem_model_set_mesh(hfem_model *md, int nnode, const double *coords, int etype, int nelem,
                                   const int *conn, const int *emat, int nmat, const double *mat, const int *eqn,
                                   int neq, int npeq)
{
    HIPFEM_REQUIRE(md && coords && conn && emat && mat && eqn, HIPFEM_EINVAL, "set_mesh: null argument");
    const int nen = nen_of(etype);
    HIPFEM_REQUIRE(nen > 0, HIPFEM_EINVAL, "set_mesh: unknown element type %d", etype);
    HIPFEM_REQUIRE(nnode > 0 && nelem >= 0 && nmat > 0 && neq >= 0 && npeq >= 0, HIPFEM_EINVAL, "set_mesh: bad sizes");
    HIPFEM_REQUIRE((int64_t)nelem * ngp_of(etype) < INT32_MAX, HIPFEM_ETOOBIG, "set_mesh: too many Gauss points");
    // host-side validation of every index the kernels will dereference
    for (int64_t i = 0; i < (int64_t)nelem * nen; i++)
        HIPFEM_REQUIRE(conn[i] >= 0 && conn[i] < nnode, HIPFEM_ERANGE, "set_mesh: conn[%lld]=%d outside [0,%d)",
                       (long long)i, conn[i], nnode);
    for (int e = 0; e < nelem; e++)
        HIPFEM_REQUIRE(emat[e] >= 0 && emat[e] < nmat, HIPFEM_ERANGE, "set_mesh: emat[%d]=%d outside [0,%d)", e, emat[e], nmat);
    for (int k = 0; k < nmat; k++) {
        const int model = (int)mat[(size_t)k * HIPFEM_MATP];
        HIPFEM_REQUIRE(model == HIPFEM_MAT_ISOLE || model == HIPFEM_MAT_MISES, HIPFEM_EINVAL,
                       "set_mesh: material %d has unknown model %d", k, model);
    }
    for (int64_t i = 0; i < (int64_t)nnode * 3; i++)
        HIPFEM_REQUIRE(eqn[i] <= neq && eqn[i] >= -npeq, HIPFEM_ERANGE, "set_mesh: eqn[%lld]=%d outside [-%d,%d]",
                       (long long)i, eqn[i], npeq, neq);
    DeviceGuard g(md->ctx->device);
    int rc = model_free(md);
    if (rc) return rc;
    md->nnode = nnode;
    md->etype = etype;
    md->nen = nen;
    md->ngp = ngp_of(etype);
    md->nelem = nelem;
    md->nmat = nmat;
    md->neq = neq;
    md->npeq = npeq;
    md->ngp_total = nelem * md->ngp;
    md->h_conn.assign(conn, conn + (size_t)nelem * nen);
    md->h_eqn.assign(eqn, eqn + (size_t)nnode * 3);
    md->has_mises = false;
    for (int k = 0; k < nmat; k++)
        if ((int)mat[(size_t)k * HIPFEM_MATP] == HIPFEM_MAT_MISES) md->has_mises = true;
    hipStream_t st = md->ctx->stream;
    HIPFEM_CHECK(hipMalloc(&md->d_coords, sizeof(double) * 3 * nnode));
    HIPFEM_CHECK(hipMalloc(&md->d_conn, sizeof(int) * std::max<size_t>(1, (size_t)nelem * nen)));
    HIPFEM_CHECK(hipMalloc(&md->d_emat, sizeof(int) * std::max(1, nelem)));
    HIPFEM_CHECK(hipMalloc(&md->d_mat, sizeof(double) * HIPFEM_MATP * nmat));
    HIPFEM_CHECK(hipMalloc(&md->d_eqn, sizeof(int) * 3 * nnode));
    HIPFEM_CHECK(hipMalloc(&md->d_bad, sizeof(int)));
    HIPFEM_CHECK(hipMalloc(&md->d_state, sizeof(double) * NST * std::max(1, md->ngp_total)));
    HIPFEM_CHECK(hipMalloc(&md->d_out3, sizeof(double) * 8));   // k_reduce3 writes 3, k_reduce6 writes 6
    HIPFEM_CHECK(hipMalloc(&md->d_uf, sizeof(double) * std::max(1, neq)));
    HIPFEM_CHECK(hipMalloc(&md->d_ff, sizeof(double) * std::max(1, neq)));
    HIPFEM_CHECK(hipMalloc(&md->d_up, sizeof(double) * std::max(1, npeq)));
    HIPFEM_CHECK(hipMalloc(&md->d_fp, sizeof(double) * std::max(1, npeq)));
    HIPFEM_CHECK(hipMemcpyAsync(md->d_coords, coords, sizeof(double) * 3 * nnode, hipMemcpyHostToDevice, st));
    if (nelem) {
        HIPFEM_CHECK(hipMemcpyAsync(md->d_conn, conn, sizeof(int) * (size_t)nelem * nen, hipMemcpyHostToDevice, st));
        HIPFEM_CHECK(hipMemcpyAsync(md->d_emat, emat, sizeof(int) * nelem, hipMemcpyHostToDevice, st));
    }
    HIPFEM_CHECK(hipMemcpyAsync(md->d_mat, mat, sizeof(double) * HIPFEM_MATP * nmat, hipMemcpyHostToDevice, st));
    HIPFEM_CHECK(hipMemcpyAsync(md->d_eqn, eqn, sizeof(int) * 3 * nnode, hipMemcpyHostToDevice, st));
    std::vector<int8_t> dofid((size_t)std::max(1, neq), 0);
    for (int64_t i = 0; i < (int64_t)nnode * 3; i++)
        if (eqn[i] > 0) dofid[eqn[i] - 1] = (int8_t)(i % 3);
    HIPFEM_CHECK(hipMalloc(&md->d_dofid, dofid.size()));
    HIPFEM_CHECK(hipMemcpyAsync(md->d_dofid, dofid.data(), dofid.size(), hipMemcpyHostToDevice, st));
    const int64_t ns = (int64_t)NST * std::max(1, md->ngp_total);
    k_fill<<<(int)std::min<int64_t>((ns + 255) / 256, 8192), 256, 0, st>>>(ns, 0.0, md->d_state);
    HIPFEM_CHECK(hipGetLastError());
    HIPFEM_CHECK(hipStreamSynchronize(st));
    return 0;
}

extern "C" int hfem_model_residual(hfem_model *md, const double *dRT, const double *dF, double *drhs, double *sums6)
{
    HIPFEM_REQUIRE(md && md->d_dofid && dF && drhs && sums6, HIPFEM_EINVAL, "residual: bad arguments");
    DeviceGuard g(md->ctx->device);
    hipStream_t st = md->ctx->stream;
    const int np = std::max(1, std::min((md->neq + 255) / 256, 2048));
    int rc = ensure_partials(md->ctx, 6 * (size_t)np);
    if (rc) return rc;
    if (md->neq) {
        k_residual_groups<<<np, 256, 0, st>>>(md->neq, md->d_dofid, dRT, dF, drhs, md->ctx->d_partials, np);
        k_reduce6<<<1, 1024, 0, st>>>(md->ctx->d_partials, np, md->d_out3);
        HIPFEM_CHECK(hipGetLastError());
        HIPFEM_CHECK(hipMemcpyAsync(md->ctx->h_pinned, md->d_out3, 6 * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPFEM_CHECK(hipStreamSynchronize(st));
        for (int k = 0; k < 6; k++) sums6[k] = md->ctx->h_pinned[k];
    } else {
        for (int k = 0; k < 6; k++) sums6[k] = 0.0;
    }
    return 0;
}

extern "C" int hfem_model_group_sums(hfem_model *md, const double *dx, const double *dy, double *sums6)
{
    HIPFEM_REQUIRE(md && md->d_dofid && dx && dy && sums6, HIPFEM_EINVAL, "group_sums: bad arguments");
    DeviceGuard g(md->ctx->device);
    hipStream_t st = md->ctx->stream;
    const int np = std::max(1, std::min((md->neq + 255) / 256, 2048));
    int rc = ensure_partials(md->ctx, 6 * (size_t)np);
    if (rc) return rc;
    if (md->neq) {
        k_group_sums<<<np, 256, 0, st>>>(md->neq, md->d_dofid, dx, dy, md->ctx->d_partials, np);
        k_reduce6<<<1, 1024, 0, st>>>(md->ctx->d_partials, np, md->d_out3);
        HIPFEM_CHECK(hipGetLastError());
        HIPFEM_CHECK(hipMemcpyAsync(md->ctx->h_pinned, md->d_out3, 6 * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPFEM_CHECK(hipStreamSynchronize(st));
        for (int k = 0; k < 6; k++) sums6[k] = md->ctx->h_pinned[k];
    } else {
        for (int k = 0; k < 6; k++) sums6[k] = 0.0;
    }
    return 0;
}

// Pattern + eoff map + coloring, built on the host from node adjacency.
extern "C" int hfem_model_build_structure(hfem_model *md, hipcsr_mtx *K)
{
    HIPFEM_REQUIRE(md && K && md->d_conn, HIPFEM_EINVAL, "build_structure: model has no mesh");
    HIPFEM_REQUIRE(K->ctx == md->ctx, HIPFEM_EINVAL, "build_structure: matrix of another context");
    DeviceGuard g(md->ctx->device);
    const int nen = md->nen, nnode = md->nnode, nelem = md->nelem, neq = md->neq;
    const int *conn = md->h_conn.data(), *eqn = md->h_eqn.data();
    // node -> elements (ascending)
    std::vector<int> nptr((size_t)nnode + 1, 0), nlst((size_t)nelem * nen);
    for (size_t i = 0; i < (size_t)nelem * nen; i++) nptr[conn[i] + 1]++;
    for (int n = 0; n < nnode; n++) nptr[n + 1] += nptr[n];
    {
        std::vector<int> fill(nptr.begin(), nptr.end() - 1);
        for (int e = 0; e < nelem; e++)
            for (int a = 0; a < nen; a++) nlst[fill[conn[(size_t)e * nen + a]]++] = e;
    }
    auto nfree = [&](int n) { return (eqn[3 * n] > 0) + (eqn[3 * n + 1] > 0) + (eqn[3 * n + 2] > 0); };
    auto firsteq = [&](int n) {
        for (int d = 0; d < 3; d++)
            if (eqn[3 * n + d] > 0) return eqn[3 * n + d];
        return 0;
    };
    // per node: sorted neighbor list (free neighbors) and row length
    std::vector<int> rowlen((size_t)neq, 0);
    std::vector<int64_t> nbptr((size_t)nnode + 1, 0);
    std::vector<int> mark((size_t)nnode, -1);
    std::vector<int> nb;
    std::vector<std::vector<int>> nbrs((size_t)nnode);
    for (int a = 0; a < nnode; a++) {
        if (!nfree(a)) continue;
        nb.clear();
        for (int k = nptr[a]; k < nptr[a + 1]; k++) {
            const int e = nlst[k];
            for (int b = 0; b < nen; b++) {
                const int nbn = conn[(size_t)e * nen + b];
                if (mark[nbn] == a || !nfree(nbn)) continue;
                mark[nbn] = a;
                nb.push_back(nbn);
            }
        }
        std::sort(nb.begin(), nb.end(), [&](int x, int y) { return firsteq(x) < firsteq(y); });
        nbrs[a] = nb;
        int len = 0;
        for (int x : nb) len += nfree(x);
        for (int d = 0; d < 3; d++)
            if (eqn[3 * a + d] > 0) rowlen[eqn[3 * a + d] - 1] = len;
    }
    std::vector<int32_t> rowptr((size_t)neq + 1, 0);
    int64_t total = 0;
    int maxlen = 0;
    for (int r = 0; r < neq; r++) {
        maxlen = std::max(maxlen, (int)rowlen[r]);
        total += rowlen[r];
        HIPFEM_REQUIRE(total < INT32_MAX, HIPFEM_ETOOBIG, "pattern exceeds int32 indices");
        rowptr[r + 1] = (int32_t)total;
    }
    std::vector<int32_t> colind((size_t)total);
    std::vector<uint16_t> eoff((size_t)nelem * nen * nen, 0xFFFF);
    std::vector<int> offmark((size_t)nnode, -1);
    for (int a = 0; a < nnode; a++) {
        if (!nfree(a)) continue;
        // column list of node a's rows and the offset of every neighbor
        std::vector<int32_t> cols;
        cols.reserve(81);
        for (int x : nbrs[a]) {
            HIPFEM_REQUIRE(cols.size() < 0xFFFF, HIPFEM_ETOOBIG, "row of node %d too long", a);
            offmark[x] = (int)cols.size();
            for (int d = 0; d < 3; d++)
                if (eqn[3 * x + d] > 0) cols.push_back(eqn[3 * x + d] - 1);
        }
        for (int d = 0; d < 3; d++)
            if (eqn[3 * a + d] > 0) std::copy(cols.begin(), cols.end(), colind.begin() + rowptr[eqn[3 * a + d] - 1]);
        for (int k = nptr[a]; k < nptr[a + 1]; k++) {
            const int e = nlst[k];
            int la = -1;
            for (int b = 0; b < nen; b++)
                if (conn[(size_t)e * nen + b] == a) { la = b; break; }
            for (int b = 0; b < nen; b++) {
                const int nbn = conn[(size_t)e * nen + b];
                if (!nfree(nbn)) continue;
                eoff[((size_t)e * nen + la) * nen + b] = (uint16_t)offmark[nbn];
            }
        }
    }
    // greedy coloring in ascending element order: smallest color not used by
    // an element sharing a node
    std::vector<uint64_t> nodemask((size_t)nnode, 0);
    std::vector<int> color((size_t)nelem, 0);
    int ncol = 0;
    for (int e = 0; e < nelem; e++) {
        uint64_t used = 0;
        for (int a = 0; a < nen; a++) used |= nodemask[conn[(size_t)e * nen + a]];
        HIPFEM_REQUIRE(~used != 0, HIPFEM_ETOOBIG, "coloring needs more than 64 colors");
        const int c = __builtin_ctzll(~used);
        color[e] = c;
        ncol = std::max(ncol, c + 1);
        for (int a = 0; a < nen; a++) nodemask[conn[(size_t)e * nen + a]] |= (1ull << c);
    }
    std::vector<int> cptr((size_t)ncol + 1, 0), clist((size_t)nelem);
    for (int e = 0; e < nelem; e++) cptr[color[e] + 1]++;
    for (int c = 0; c < ncol; c++) cptr[c + 1] += cptr[c];
    {
        std::vector<int> fill(cptr.begin(), cptr.end() - 1);
        for (int e = 0; e < nelem; e++) clist[fill[color[e]]++] = e;
    }
    int rc = csr_alloc(K, neq, rowptr, colind);
    if (rc) return rc;
    hipStream_t st = md->ctx->stream;
    if (md->d_eoff) HIPFEM_CHECK(hipFree(md->d_eoff));
    if (md->d_colors) HIPFEM_CHECK(hipFree(md->d_colors));
    md->d_eoff = nullptr;
    md->d_colors = nullptr;
    HIPFEM_CHECK(hipMalloc(&md->d_eoff, sizeof(uint16_t) * std::max<size_t>(1, eoff.size())));
    HIPFEM_CHECK(hipMalloc(&md->d_colors, sizeof(int) * std::max(1, nelem)));
    if (!eoff.empty())
        HIPFEM_CHECK(hipMemcpyAsync(md->d_eoff, eoff.data(), sizeof(uint16_t) * eoff.size(), hipMemcpyHostToDevice, st));
    if (nelem) HIPFEM_CHECK(hipMemcpyAsync(md->d_colors, clist.data(), sizeof(int) * nelem, hipMemcpyHostToDevice, st));
    HIPFEM_CHECK(hipStreamSynchronize(st));
    md->color_ptr = cptr;
    md->ncolors = ncol;
    md->max_rowlen = maxlen;
    md->h_ecolor.assign(color.begin(), color.end());
    md->h_nodemask = nodemask;
    md->pattern_id = K->pattern_id;
    return 0;
}

static int check_bad(hfem_model *md, const char *what)
{
    int bad = INT32_MAX;
    HIPFEM_CHECK(hipMemcpyAsync(&bad, md->d_bad, sizeof(int), hipMemcpyDeviceToHost, md->ctx->stream));
    HIPFEM_CHECK(hipStreamSynchronize(md->ctx->stream));
    // FEI3dTetLin::evaldNdx aborts with "negative volume" (fei3dtetlin.C:146)
    HIPFEM_REQUIRE(bad == INT32_MAX, HIPFEM_ENUMERIC, "%s: degenerate or inverted element %d", what, bad + 1);
    return 0;
}

static int reset_bad(hfem_model *md)
{
    const int big = INT32_MAX;
    HIPFEM_CHECK(hipMemcpyAsync(md->d_bad, &big, sizeof(int), hipMemcpyHostToDevice, md->ctx->stream));
    return 0;
}

static int node_adjacency(hfem_model *md);

extern "C" int hfem_model_assemble_stiffness(hfem_model *md, hipcsr_mtx *K, int mode)
{
    HIPFEM_REQUIRE(md && K && K->has_structure, HIPFEM_EINVAL, "assemble: no structure");
    HIPFEM_REQUIRE(md->pattern_id == K->pattern_id, HIPFEM_EINVAL,
                   "assemble: matrix pattern was not built by hfem_model_build_structure for this model");
    HIPFEM_REQUIRE(mode == 0 || mode == 1, HIPFEM_EINVAL, "assemble: mode must be 0 or 1");
    DeviceGuard g(md->ctx->device);
    hipStream_t st = md->ctx->stream;
    int rc = reset_bad(md);
    if (rc) return rc;
    // LSpace by rows: every element's blocks stored in one launch, then each
    // node's rows summed in ascending element order and written once (no
    // read-modify-write of partial lines); else one launch per color adding
    // into the zeroed values
    bool by_rows = md->etype == HIPFEM_LSPACE && md->nelem > 0 && md->max_rowlen <= ASM_LMAX && !md->kbuf_failed;
    if (by_rows) {
        rc = node_adjacency(md);
        if (rc) return rc;
        if (!md->d_kbuf && hipMalloc(&md->d_kbuf, sizeof(double) * 64 * 9 * (size_t)md->nelem) != hipSuccess) {
            (void)hipGetLastError();   // the colored form needs no buffer
            md->d_kbuf = nullptr;
            md->kbuf_failed = true;
            by_rows = false;
        }
    }
    if (by_rows) {
        k_assemble<HIPFEM_LSPACE><<<md->nelem, 64, 0, st>>>(nullptr, md->nelem, md->d_coords, md->d_conn, md->d_eqn,
                                                            md->d_emat, md->d_mat, md->d_eoff, md->d_state,
                                                            md->ngp_total, mode, K->d_rowptr, K->d_val, md->d_bad,
                                                            md->sur ? md->d_sflag : nullptr, md->d_dsur, md->d_kbuf);
        HIPFEM_CHECK(hipGetLastError());
        k_assemble_gather<8><<<md->nnode, 64, 0, st>>>(md->nnode, md->d_fnptr, md->d_fnadj, md->d_conn, md->d_eqn,
                                                       md->d_eoff, md->d_kbuf, K->d_rowptr, K->d_val);
        HIPFEM_CHECK(hipGetLastError());
        K->h_stage.clear();   // every value was written: what hipcsr_zero resets
        K->staged = false;
    } else {
        rc = hipcsr_zero(K);
        if (rc) return rc;
    }
    for (int c = 0; c < md->ncolors && !by_rows; c++) {
        const int n = md->color_ptr[c + 1] - md->color_ptr[c];
        if (!n) continue;
        const int *el = md->d_colors + md->color_ptr[c];
#define LAUNCH_ASM(ET)                                                                                              \
    k_assemble<ET><<<n, 64, 0, st>>>(el, n, md->d_coords, md->d_conn, md->d_eqn, md->d_emat, md->d_mat, md->d_eoff, \
                                     md->d_state, md->ngp_total, mode, K->d_rowptr, K->d_val, md->d_bad,     \
                                     md->sur ? md->d_sflag : nullptr, md->d_dsur, nullptr)
        if (md->etype == HIPFEM_LSPACE) LAUNCH_ASM(HIPFEM_LSPACE);
        else if (md->etype == HIPFEM_LTRSPACE) LAUNCH_ASM(HIPFEM_LTRSPACE);
        else LAUNCH_ASM(HIPFEM_QSPACE);
#undef LAUNCH_ASM
        HIPFEM_CHECK(hipGetLastError());
    }
    K->version++;
    rc = check_bad(md, "assemble");
    if (rc) return rc;
    // element-by-element view of what was just assembled (LSpace): the FP32
    // tangent descriptors of this state and a link from K to the model
    if (md->etype == HIPFEM_LSPACE && md->ngp_total > 0) {
        md->tdesc_used = md->has_mises && mode == 1;
        if (md->tdesc_used) {
            if (!md->d_tdesc) HIPFEM_CHECK(hipMalloc(&md->d_tdesc, sizeof(float) * 6 * (size_t)md->ngp_total));
            k_tangent_desc<<<(md->ngp_total + 255) / 256, 256, 0, st>>>(md->ngp_total, md->ngp, md->d_emat,
                                                                          md->d_mat, md->d_state, mode, md->d_tdesc,
                                                                          md->sur ? md->d_sflag : nullptr,
                                                                          md->d_dsur, md->d_dsnap);
            HIPFEM_CHECK(hipGetLastError());
        }
        if (!md->d_erel && md->nelem) {
            // per-element coefficients of the trilinear coordinate field, in
            // FP64 rounded once to FP32 (k_mf_hex would otherwise transform
            // FP32 coordinates each time)
            HIPFEM_CHECK(hipMalloc(&md->d_erel, sizeof(float) * 21 * (size_t)md->nelem));
            k_mf_erel<<<(md->nelem + 255) / 256, 256, 0, st>>>(md->nelem, md->d_coords, md->d_conn, md->d_erel);
            HIPFEM_CHECK(hipGetLastError());
        }
        if (!md->eloc_checked && md->nelem) {
            // k_mf_hex's equation words, where each node's free dofs carry
            // consecutive equation numbers in dof order (the node-by-node
            // numbering); otherwise the kernel reads conn and eqn
            md->eloc_checked = true;
            bool ok = md->neq < (1 << 29);
            std::vector<uint32_t> nodew(ok ? (size_t)md->nnode : 0);
            for (int n = 0; ok && n < md->nnode; n++) {
                uint32_t fm = 0;
                int first = 0, k = 0;
                for (int d = 0; d < 3 && ok; d++) {
                    const int q = md->h_eqn[3 * (size_t)n + d];
                    if (q <= 0) continue;
                    if (!fm)
                        first = q - 1;
                    else if (q - 1 != first + k)
                        ok = false;
                    fm |= 1u << d;
                    k++;
                }
                nodew[n] = (uint32_t)first | fm << 29;
            }
            if (ok) {
                std::vector<uint32_t> eloc((size_t)md->nelem * 8);
                for (size_t e = 0; e < (size_t)md->nelem; e++)
                    for (int g = 0; g < 8; g++)
                        eloc[e * 8 + g] = nodew[md->h_conn[e * 8 + ((0x26371504 >> (4 * g)) & 15)]];
                HIPFEM_CHECK(hipMalloc(&md->d_eloc, sizeof(uint32_t) * eloc.size()));
                HIPFEM_CHECK(hipMemcpy(md->d_eloc, eloc.data(), sizeof(uint32_t) * eloc.size(), hipMemcpyHostToDevice));
            }
        }
        if (!md->d_ecolor && md->ncolors <= 16 && md->nelem && md->neq) {
            // slot form: element colors and the colors touching each free equation
            std::vector<uint16_t> eqmask((size_t)md->neq, 0);
            for (int n = 0; n < md->nnode; n++)
                for (int d = 0; d < 3; d++) {
                    const int q = md->h_eqn[3 * (size_t)n + d];
                    if (q > 0) eqmask[q - 1] = (uint16_t)md->h_nodemask[n];
                }
            HIPFEM_CHECK(hipMalloc(&md->d_ecolor, (size_t)md->nelem));
            HIPFEM_CHECK(hipMalloc(&md->d_eqmask, sizeof(uint16_t) * (size_t)md->neq));
            HIPFEM_CHECK(hipMalloc(&md->d_yslot, sizeof(float) * (size_t)md->ncolors * md->neq));
            // the slot of a color that touches no element at an equation is
            // never written and stays zero, so a sum may read every slot
            const int64_t ns = (int64_t)md->ncolors * md->neq;
            k_zero_f32<<<(int)std::min<int64_t>((ns + 255) / 256, 8192), 256, 0, st>>>(ns, md->d_yslot);
            HIPFEM_CHECK(hipGetLastError());
            HIPFEM_CHECK(hipMemcpy(md->d_ecolor, md->h_ecolor.data(), (size_t)md->nelem, hipMemcpyHostToDevice));
            HIPFEM_CHECK(hipMemcpy(md->d_eqmask, eqmask.data(), sizeof(uint16_t) * eqmask.size(), hipMemcpyHostToDevice));
        }
        if (!md->d_nptr) {
            // node -> element adjacency of the gather form (built once per mesh)
            std::vector<int> cnt((size_t)md->nnode + 1, 0), adj((size_t)md->nelem * 8);
            for (size_t i = 0; i < (size_t)md->nelem * 8; i++) cnt[md->h_conn[i] + 1]++;
            for (int n = 0; n < md->nnode; n++) cnt[n + 1] += cnt[n];
            std::vector<int> fill(cnt.begin(), cnt.end() - 1);
            for (int e = 0; e < md->nelem; e++)
                for (int a = 0; a < 8; a++) adj[fill[md->h_conn[(size_t)e * 8 + a]]++] = e * 8 + a;
            HIPFEM_CHECK(hipMalloc(&md->d_nptr, sizeof(int) * cnt.size()));
            HIPFEM_CHECK(hipMalloc(&md->d_nadj, sizeof(int) * std::max<size_t>(1, adj.size())));
            HIPFEM_CHECK(hipMalloc(&md->d_ebuf, sizeof(float) * 24 * std::max<size_t>(1, (size_t)md->nelem)));
            HIPFEM_CHECK(hipMemcpy(md->d_nptr, cnt.data(), sizeof(int) * cnt.size(), hipMemcpyHostToDevice));
            if (!adj.empty())
                HIPFEM_CHECK(hipMemcpy(md->d_nadj, adj.data(), sizeof(int) * adj.size(), hipMemcpyHostToDevice));
        }
        if (K->mf_model && K->mf_model != md) model_unlink(K->mf_model, K);
        // the descriptors now describe this assembly only
        const std::vector<hipcsr_mtx *> others = md->linked;
        for (hipcsr_mtx *k : others)
            if (k != K) model_unlink(md, k);
        if (md->linked.empty()) md->linked.push_back(K);
        K->mf_model = md;
        K->mf_version = K->version;
    }
    return 0;
}

namespace hipfem {
bool model_mf_available(const hipcsr_mtx *m)
{
    return m->mf_enabled && m->mf_model && m->mf_version == m->version && m->mf_model->pattern_id == m->pattern_id;
}

void model_unlink(hfem_model *md, hipcsr_mtx *m)
{
    md->linked.erase(std::remove(md->linked.begin(), md->linked.end(), m), md->linked.end());
    m->mf_model = nullptr;
    m->mf_version = -1;
}

// k_mf_hex over nel elements (elems: one color's list, or NULL for all),
// in the lookup and tangent forms of this model and its last assembly
static void launch_mf_hex(const hfem_model *md, hipStream_t s, const int *elems, int nel, const float *x, double *y,
                          float *ebuf, float *yslot, const PcgState *st)
{
    const float *desc = md->tdesc_used ? md->d_tdesc : nullptr;
    const int grid = (nel + MF_EPB - 1) / MF_EPB;
    const uint8_t *ecolor = yslot ? md->d_ecolor : nullptr;
    const int neq = yslot ? md->neq : 0;
#define MF_ARGS                                                                                                    \
    elems, nel, md->d_erel, md->d_conn, md->d_eqn, md->d_eloc, md->d_emat, md->d_mat, desc, md->d_dsnap, md->ngp_total, \
        x, y, ebuf, yslot, ecolor, neq, st
    if (md->d_eloc && desc)
        k_mf_hex<true, true><<<grid, 256, 0, s>>>(MF_ARGS);
    else if (md->d_eloc)
        k_mf_hex<true, false><<<grid, 256, 0, s>>>(MF_ARGS);
    else if (desc)
        k_mf_hex<false, true><<<grid, 256, 0, s>>>(MF_ARGS);
    else
        k_mf_hex<false, false><<<grid, 256, 0, s>>>(MF_ARGS);
#undef MF_ARGS
}

// the element pass of the slot form: one launch over every element
static int mf_slot_pass(hfem_model *md, const float *x, const PcgState *st)
{
    if (!md->nelem || !md->neq) return 0;
    launch_mf_hex(md, md->ctx->stream, nullptr, md->nelem, x, nullptr, nullptr, md->d_yslot, st);
    HIPFEM_CHECK(hipGetLastError());
    return 0;
}

int model_mf_slots_f32(hfem_model *md, const float *x, const PcgState *st, const float **yslot,
                       const uint16_t **eqmask, int *ncolors)
{
    *yslot = nullptr;
    *eqmask = nullptr;
    *ncolors = 0;
    if (!(md->mf_form == 2 && md->d_yslot && md->d_erel)) return 0;
    int rc = mf_slot_pass(md, x, st);
    if (rc) return rc;
    *yslot = md->d_yslot;
    *eqmask = md->d_eqmask;
    *ncolors = md->ncolors;
    return 0;
}

int model_mf_apply_f32(hfem_model *md, const float *x, double *y, const PcgState *st)
{
    hipStream_t s = md->ctx->stream;
    if (md->mf_form == 2 && md->d_yslot && md->d_erel) {
        // slot form: every element in one launch, each writing its forces
        // into its color's FP32 slot (no element of a color shares a node),
        // then one FP64 sum per equation over the colors in ascending order
        // (the order of the per-color launches: the same result, bit for bit)
        int rc = mf_slot_pass(md, x, st);
        if (rc) return rc;
        k_mf_slot_sum<<<std::min(8192, (md->neq + 255) / 256), 256, 0, s>>>(md->neq, md->ncolors, md->d_eqmask,
                                                                            md->d_yslot, y, st);
        HIPFEM_CHECK(hipGetLastError());
        return 0;
    }
    if (md->d_nptr && md->mf_form == 1) {
        // gather form: every element in one launch, then one FP64 sum per node
        if (md->nelem && md->neq) launch_mf_hex(md, s, nullptr, md->nelem, x, y, md->d_ebuf, nullptr, st);
        k_mf_gather<<<std::min(8192, (md->nnode + 255) / 256), 256, 0, s>>>(md->nnode, md->d_nptr, md->d_nadj,
                                                                              md->d_eqn, md->d_ebuf, y, st);
        HIPFEM_CHECK(hipGetLastError());
        return 0;
    }
    if (md->neq) k_fill_d<<<std::min(4096, (md->neq + 255) / 256), 256, 0, s>>>(md->neq, y, st);
    for (int c = 0; c < md->ncolors; c++) {
        const int n = md->color_ptr[c + 1] - md->color_ptr[c];
        if (!n || !md->neq) continue;
        launch_mf_hex(md, s, md->d_colors + md->color_ptr[c], n, x, y, nullptr, nullptr, st);
    }
    HIPFEM_CHECK(hipGetLastError());
    return 0;
}
}  // namespace hipfem

// node -> (element * nen + local node) lists in ascending element order,
// built once per mesh (internal-force gather, stiffness assembly by rows)
static int node_adjacency(hfem_model *md)
{
    if (md->d_fnptr || !md->nelem) return 0;
    const int nen = md->nen;
    std::vector<int> cnt((size_t)md->nnode + 1, 0), adj((size_t)md->nelem * nen);
    for (size_t i = 0; i < (size_t)md->nelem * nen; i++) cnt[md->h_conn[i] + 1]++;
    for (int n = 0; n < md->nnode; n++) cnt[n + 1] += cnt[n];
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (int e = 0; e < md->nelem; e++)
        for (int a = 0; a < nen; a++) adj[fill[md->h_conn[(size_t)e * nen + a]]++] = e * nen + a;
    HIPFEM_CHECK(hipMalloc(&md->d_fnptr, sizeof(int) * cnt.size()));
    HIPFEM_CHECK(hipMalloc(&md->d_fnadj, sizeof(int) * adj.size()));
    HIPFEM_CHECK(hipMemcpy(md->d_fnptr, cnt.data(), sizeof(int) * cnt.size(), hipMemcpyHostToDevice));
    HIPFEM_CHECK(hipMemcpy(md->d_fnadj, adj.data(), sizeof(int) * adj.size(), hipMemcpyHostToDevice));
    return 0;
}

static int internal_dev(hfem_model *md, const double *uf, const double *up, double *ff, double *fp, double *ebe3,
                        int update, int op = 0)
{
    hipStream_t st = md->ctx->stream;
    HIPFEM_REQUIRE(md->ncolors > 0 || md->nelem == 0, HIPFEM_EINVAL,
                   "internal forces: call hfem_model_build_structure first (it builds the element coloring)");
    int epb = md->etype == HIPFEM_LSPACE ? IntCfg<HIPFEM_LSPACE>::EPB
              : md->etype == HIPFEM_LTRSPACE ? IntCfg<HIPFEM_LTRSPACE>::EPB : IntCfg<HIPFEM_QSPACE>::EPB;
    const int npart = (md->nelem + epb - 1) / epb;
    int rc = ensure_partials(md->ctx, 3 * (size_t)std::max(1, npart));
    if (rc) return rc;
    rc = node_adjacency(md);
    if (rc) return rc;
    if (!md->d_fbuf && md->nelem)
        HIPFEM_CHECK(hipMalloc(&md->d_fbuf, sizeof(double) * 3 * md->nen * (size_t)md->nelem));
    rc = reset_bad(md);
    if (rc) return rc;
    if (ff && md->neq) k_fill<<<std::min(8192, (md->neq + 255) / 256), 256, 0, st>>>(md->neq, 0.0, ff);
    if (fp && md->npeq) k_fill<<<std::min(8192, (md->npeq + 255) / 256), 256, 0, st>>>(md->npeq, 0.0, fp);
    double *part = md->ctx->d_partials;
    SurArgs sur;
    if (md->sur && op == 0 && md->has_mises && md->ngp_total > 0) {
        // queue the points beyond the band, infer them in one BF16 batch
        HIPFEM_CHECK(hipMemsetAsync(md->d_qcount, 0, sizeof(int), st));
        const int g2 = (md->ngp_total + 255) / 256;
#define LAUNCH_CLS(ET)                                                                                           \
    k_sur_classify<ET><<<g2, 256, 0, st>>>(md->nelem, md->d_coords, md->d_conn, md->d_eqn, md->d_emat, md->d_mat, \
                                           uf, up, md->d_state, md->ngp_total, md->sur_band, md->sur_scales[0],   \
                                           md->d_srej, md->d_slot, md->d_qcount, md->d_sx)
        if (md->etype == HIPFEM_LSPACE) LAUNCH_CLS(HIPFEM_LSPACE);
        else if (md->etype == HIPFEM_LTRSPACE) LAUNCH_CLS(HIPFEM_LTRSPACE);
        else LAUNCH_CLS(HIPFEM_QSPACE);
#undef LAUNCH_CLS
        HIPFEM_CHECK(hipGetLastError());
        int nq = 0;
        HIPFEM_CHECK(hipMemcpyAsync(md->ctx->h_pinned, md->d_qcount, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPFEM_CHECK(hipStreamSynchronize(st));
        memcpy(&nq, md->ctx->h_pinned, sizeof(int));
        HIPFEM_REQUIRE(nq >= 0 && nq <= md->ngp_total, HIPFEM_EHIP, "surrogate queue count %d out of range", nq);
        md->queued_total += (unsigned long long)nq;
        rc = mlp_infer_bf16_dev(md->sur, nq, md->d_sx, md->d_sy);
        if (rc) return rc;
        sur.slot = md->d_slot;
        sur.y = md->d_sy;
        sur.dsur = md->d_dsur;
        sur.flag = md->d_sflag;
        sur.rejected = md->d_srej;
        sur.counts = md->d_scounts;
        sur.sig_scale = md->sur_scales[1];
        sur.tan_scale = md->sur_scales[2];
        sur.tol = md->sur_tol;
    }
    if (md->nelem) {
        // every element in one launch (the Gauss-point state is read and
        // written in whole lines), each storing its force vector; then one
        // FP64 sum per node in the fixed order of its element list
#define LAUNCH_INT(ET)                                                                                              \
    k_internal<ET><<<npart, 64, 0, st>>>(nullptr, md->nelem, md->d_coords, md->d_conn, md->d_eqn, md->d_emat,        \
                                         md->d_mat, uf, up, md->d_state, md->ngp_total, update, nullptr, nullptr,    \
                                         md->d_fbuf, part, npart, md->d_bad, op, sur)
        if (md->etype == HIPFEM_LSPACE) LAUNCH_INT(HIPFEM_LSPACE);
        else if (md->etype == HIPFEM_LTRSPACE) LAUNCH_INT(HIPFEM_LTRSPACE);
        else LAUNCH_INT(HIPFEM_QSPACE);
#undef LAUNCH_INT
        HIPFEM_CHECK(hipGetLastError());
        if (ff || fp) {
            k_gather_nodal<<<std::min(8192, (md->nnode + 255) / 256), 256, 0, st>>>(
                md->nnode, md->nen, md->d_fnptr, md->d_fnadj, md->d_eqn, md->d_fbuf, ff, fp);
            HIPFEM_CHECK(hipGetLastError());
        }
    }
    if (ebe3) {
        k_reduce3<<<1, 1024, 0, st>>>(part, npart, md->d_out3);
        HIPFEM_CHECK(hipGetLastError());
        HIPFEM_CHECK(hipMemcpyAsync(ebe3, md->d_out3, 3 * sizeof(double), hipMemcpyDeviceToHost, st));
    }
    return check_bad(md, "internal forces");
}

extern "C" int hfem_model_internal_forces_dev(hfem_model *md, const double *du_free, const double *du_presc,
                                              double *df_free, double *df_presc, double *ebe3, int update_state)
{
    HIPFEM_REQUIRE(md && md->d_conn && (du_free || md->neq == 0), HIPFEM_EINVAL, "internal forces: bad arguments");
    DeviceGuard g(md->ctx->device);
    return internal_dev(md, du_free, du_presc, df_free, df_presc, ebe3, update_state);
}

extern "C" int hfem_model_internal_forces(hfem_model *md, const double *u_free, const double *u_presc, double *f_free,
                                          double *f_presc, double *ebe3, int update_state)
{
    HIPFEM_REQUIRE(md && md->d_conn && (u_free || md->neq == 0), HIPFEM_EINVAL, "internal forces: bad arguments");
    DeviceGuard g(md->ctx->device);
    hipStream_t st = md->ctx->stream;
    if (md->neq) HIPFEM_CHECK(hipMemcpyAsync(md->d_uf, u_free, sizeof(double) * md->neq, hipMemcpyHostToDevice, st));
    if (u_presc && md->npeq)
        HIPFEM_CHECK(hipMemcpyAsync(md->d_up, u_presc, sizeof(double) * md->npeq, hipMemcpyHostToDevice, st));
    int rc = internal_dev(md, md->d_uf, (u_presc && md->npeq) ? md->d_up : nullptr, md->d_ff,
                          md->npeq ? md->d_fp : nullptr, ebe3, update_state);
    if (rc) return rc;
    if (f_free && md->neq) HIPFEM_CHECK(hipMemcpyAsync(f_free, md->d_ff, sizeof(double) * md->neq, hipMemcpyDeviceToHost, st));
    if (f_presc && md->npeq)
        HIPFEM_CHECK(hipMemcpyAsync(f_presc, md->d_fp, sizeof(double) * md->npeq, hipMemcpyDeviceToHost, st));
    HIPFEM_CHECK(hipStreamSynchronize(st));
    return 0;
}

extern "C" int hfem_model_set_surrogate(hfem_model *md, hfem_mlp *net, double band, double tol)
{
    HIPFEM_REQUIRE(md && md->d_conn, HIPFEM_EINVAL, "set_surrogate: no mesh");
    DeviceGuard g(md->ctx->device);
    if (!net) {
        md->sur = nullptr;
        return 0;
    }
    HIPFEM_REQUIRE(band >= 0.0 && tol > 0.0, HIPFEM_EINVAL, "set_surrogate: band must be >= 0 and tol > 0");
    hipfem_ctx *nc = nullptr;
    double m4[4];
    float sc[3];
    mlp_info(net, &nc, m4, sc);
    HIPFEM_REQUIRE(nc == md->ctx, HIPFEM_EINVAL, "set_surrogate: network and model belong to different contexts");
    // every MisesMat row of the mesh must be the material the network learned
    std::vector<double> mat((size_t)HIPFEM_MATP * md->nmat);
    HIPFEM_CHECK(hipStreamSynchronize(md->ctx->stream));
    HIPFEM_CHECK(hipMemcpy(mat.data(), md->d_mat, sizeof(double) * mat.size(), hipMemcpyDeviceToHost));
    for (int k = 0; k < md->nmat; k++) {
        const double *r = mat.data() + (size_t)HIPFEM_MATP * k;
        if ((int)r[0] != HIPFEM_MAT_MISES) continue;
        for (int i = 0; i < 4; i++)
            HIPFEM_REQUIRE(std::fabs(r[1 + i] - m4[i]) <= 1e-12 * std::max(1.0, std::fabs(m4[i])), HIPFEM_EINVAL,
                           "set_surrogate: material %d differs from the network's (E nu sig0 H)", k);
    }
    const size_t n = (size_t)std::max(1, md->ngp_total);
    if (!md->d_slot) {
        HIPFEM_CHECK(hipMalloc(&md->d_slot, sizeof(int) * n));
        HIPFEM_CHECK(hipMalloc(&md->d_qcount, sizeof(int)));
        HIPFEM_CHECK(hipMalloc(&md->d_sx, sizeof(float) * n * 13));
        HIPFEM_CHECK(hipMalloc(&md->d_sy, sizeof(float) * n * 27));
        HIPFEM_CHECK(hipMalloc(&md->d_dsur, sizeof(float) * n * 21));
        HIPFEM_CHECK(hipMalloc(&md->d_dsnap, sizeof(float) * n * 21));
        HIPFEM_CHECK(hipMalloc(&md->d_sflag, n));
        HIPFEM_CHECK(hipMalloc(&md->d_srej, n));
        HIPFEM_CHECK(hipMalloc(&md->d_scounts, 2 * sizeof(unsigned long long)));
        HIPFEM_CHECK(hipMemset(md->d_sflag, 0, n));
        HIPFEM_CHECK(hipMemset(md->d_srej, 0, n));
        HIPFEM_CHECK(hipMemset(md->d_scounts, 0, 2 * sizeof(unsigned long long)));
    }
    md->sur = net;
    md->sur_band = band;
    md->sur_tol = tol;
    for (int i = 0; i < 3; i++) md->sur_scales[i] = sc[i];
    return 0;
}

extern "C" int hfem_model_set_element_operator(hfem_model *md, int form)
{
    HIPFEM_REQUIRE(md && form >= 0 && form <= 2, HIPFEM_EINVAL, "element operator form must be 0, 1 or 2");
    md->mf_form = form;
    return 0;
}

extern "C" int hfem_model_surrogate_stats(hfem_model *md, int64_t *counts3, int reset)
{
    HIPFEM_REQUIRE(md && counts3, HIPFEM_EINVAL, "surrogate_stats: null argument");
    DeviceGuard g(md->ctx->device);
    unsigned long long c[2] = {0, 0};
    HIPFEM_CHECK(hipStreamSynchronize(md->ctx->stream));
    if (md->d_scounts) HIPFEM_CHECK(hipMemcpy(c, md->d_scounts, sizeof c, hipMemcpyDeviceToHost));
    counts3[0] = (int64_t)md->queued_total;
    counts3[1] = (int64_t)c[0];
    counts3[2] = (int64_t)c[1];
    if (reset) {
        md->queued_total = 0;
        if (md->d_scounts) HIPFEM_CHECK(hipMemset(md->d_scounts, 0, sizeof c));
    }
    return 0;
}

extern "C" int hfem_model_tangent_product_dev(hfem_model *md, int mode, const double *dx_free, const double *dx_presc,
                                              double *dy_free, double *dy_presc)
{
    HIPFEM_REQUIRE(md && md->d_conn, HIPFEM_EINVAL, "tangent product: no mesh");
    HIPFEM_REQUIRE(mode == 0 || mode == 1, HIPFEM_EINVAL, "tangent product: mode must be 0 or 1");
    HIPFEM_REQUIRE(dy_free || dy_presc, HIPFEM_EINVAL, "tangent product: no output");
    DeviceGuard g(md->ctx->device);
    return internal_dev(md, dx_free, md->npeq ? dx_presc : nullptr, dy_free, md->npeq ? dy_presc : nullptr, nullptr,
                        0, mode + 1);
}

extern "C" int hfem_model_commit(hfem_model *md)
{
    HIPFEM_REQUIRE(md && md->d_state, HIPFEM_EINVAL, "commit: no mesh");
    DeviceGuard g(md->ctx->device);
    if (md->ngp_total) {
        k_commit<<<(md->ngp_total + 255) / 256, 256, 0, md->ctx->stream>>>(md->ngp_total, md->d_state);
        HIPFEM_CHECK(hipGetLastError());
        // the committed state's tangent is MisesMat's own (temp = committed)
        if (md->d_sflag) HIPFEM_CHECK(hipMemsetAsync(md->d_sflag, 0, (size_t)md->ngp_total, md->ctx->stream));
        if (md->d_srej) HIPFEM_CHECK(hipMemsetAsync(md->d_srej, 0, (size_t)md->ngp_total, md->ctx->stream));
    }
    return 0;
}

extern "C" int hfem_model_restart(hfem_model *md)
{
    HIPFEM_REQUIRE(md && md->d_state, HIPFEM_EINVAL, "restart: no mesh");
    DeviceGuard g(md->ctx->device);
    if (md->ngp_total) {
        k_restart<<<(md->ngp_total + 255) / 256, 256, 0, md->ctx->stream>>>(md->ngp_total, md->d_state);
        HIPFEM_CHECK(hipGetLastError());
        if (md->d_sflag) HIPFEM_CHECK(hipMemsetAsync(md->d_sflag, 0, (size_t)md->ngp_total, md->ctx->stream));
        if (md->d_srej) HIPFEM_CHECK(hipMemsetAsync(md->d_srej, 0, (size_t)md->ngp_total, md->ctx->stream));
    }
    return 0;
}

extern "C" int hfem_model_reset_state(hfem_model *md)
{
    HIPFEM_REQUIRE(md && md->d_state, HIPFEM_EINVAL, "reset: no mesh");
    DeviceGuard g(md->ctx->device);
    const int64_t ns = (int64_t)NST * std::max(1, md->ngp_total);
    k_fill<<<(int)std::min<int64_t>((ns + 255) / 256, 8192), 256, 0, md->ctx->stream>>>(ns, 0.0, md->d_state);
    HIPFEM_CHECK(hipGetLastError());
    if (md->d_sflag) HIPFEM_CHECK(hipMemsetAsync(md->d_sflag, 0, (size_t)std::max(1, md->ngp_total), md->ctx->stream));
    if (md->d_srej) HIPFEM_CHECK(hipMemsetAsync(md->d_srej, 0, (size_t)std::max(1, md->ngp_total), md->ctx->stream));
    return 0;
}

extern "C" int hfem_model_get_state(hfem_model *md, double *state)
{
    HIPFEM_REQUIRE(md && md->d_state && state, HIPFEM_EINVAL, "get_state: bad arguments");
    DeviceGuard g(md->ctx->device);
    const size_t n = md->ngp_total;
    std::vector<double> soa(NST * std::max<size_t>(1, n));
    HIPFEM_CHECK(hipStreamSynchronize(md->ctx->stream));
    HIPFEM_CHECK(hipMemcpy(soa.data(), md->d_state, sizeof(double) * NST * n, hipMemcpyDeviceToHost));
    for (size_t g2 = 0; g2 < n; g2++)
        for (int k = 0; k < NST; k++) state[g2 * NST + k] = soa[k * n + g2];
    return 0;
}

extern "C" int hfem_model_set_state(hfem_model *md, const double *state)
{
    HIPFEM_REQUIRE(md && md->d_state && state, HIPFEM_EINVAL, "set_state: bad arguments");
    DeviceGuard g(md->ctx->device);
    const size_t n = md->ngp_total;
    std::vector<double> soa(NST * std::max<size_t>(1, n));
    for (size_t g2 = 0; g2 < n; g2++)
        for (int k = 0; k < NST; k++) soa[k * n + g2] = state[g2 * NST + k];
    HIPFEM_CHECK(hipStreamSynchronize(md->ctx->stream));
    HIPFEM_CHECK(hipMemcpy(md->d_state, soa.data(), sizeof(double) * NST * n, hipMemcpyHostToDevice));
    return 0;
}

extern "C" int hfem_model_ngp_total(const hfem_model *md) { return md ? md->ngp_total : 0; }
extern "C" int hfem_model_ncolors(const hfem_model *md) { return md ? md->ncolors : 0; }

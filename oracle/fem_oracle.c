/*
 * fem_oracle.c -- CPU restatement of OOFEM's assembly-and-solve hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  Loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg only.
 *
 * Each routine follows the arithmetic order of the reference routine it
 * names, so that element matrices agree with OOFEM's to the last bits where
 * the reference's own order is fixed (non-LAPACK FloatMatrix build, the
 * default).  Reference paths are relative to the OOFEM source root.
 */
#include "fem_oracle.h"

#ifdef _OPENMP
#include <omp.h>
#endif

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* Gauss rules: gaussintegrationrule.C:190-216 (SetUpPointsOnCube, loop
 * order xi -> eta -> zeta), line constants :1436-1463 (15-digit literals),
 * tetrahedron 1-point rule :492-502.                                    */
/* ------------------------------------------------------------------ */
static const double GL2_X[2] = {-0.577350269189626, 0.577350269189626};
static const double GL2_W[2] = {1.0, 1.0};
static const double GL3_X[3] = {-0.774596669241483, 0.0, 0.774596669241483};
static const double GL3_W[3] = {0.555555555555555, 0.888888888888888, 0.555555555555555};

int orc_nen(int etype)
{
    switch (etype) {
    case ORC_LSPACE: return 8;     /* lspace.C:70 numberOfDofMans */
    case ORC_LTRSPACE: return 4;
    case ORC_QSPACE: return 20;    /* qspace.C:56 */
    }
    return 0;
}

int orc_ngp(int etype)
{
    switch (etype) {
    case ORC_LSPACE: return 8;     /* lspace.C:71 */
    case ORC_LTRSPACE: return 1;   /* ltrspace.C:69 */
    case ORC_QSPACE: return 27;    /* qspace.C:57 */
    }
    return 0;
}

int orc_gauss(int etype, double *xi, double *w)
{
    if (etype == ORC_LTRSPACE) {
        xi[0] = xi[1] = xi[2] = 1.0 / 4.0;
        w[0] = 1.0 / 6.0;
        return 1;
    }
    const double *X = (etype == ORC_QSPACE) ? GL3_X : GL2_X;
    const double *W = (etype == ORC_QSPACE) ? GL3_W : GL2_W;
    int n = (etype == ORC_QSPACE) ? 3 : 2, c = 0;
    if (etype != ORC_LSPACE && etype != ORC_QSPACE) return 0;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
            for (int k = 0; k < n; k++) {
                xi[3 * c + 0] = X[i];
                xi[3 * c + 1] = X[j];
                xi[3 * c + 2] = X[k];
                w[c] = W[i] * W[j] * W[k];
                c++;
            }
    return c;
}

/* ------------------------------------------------------------------ */
/* dN/dxi: FEI3dHexaLin::evaldNdxi fei3dhexalin.C:127-175 (8 nodes; nodes
 * 1-4 on zeta=+1, 5-8 on zeta=-1), FEI3dHexaQuad::evaldNdxi
 * fei3dhexaquad.C:191-282 (20-node serendipity).  dN is nen x 3.      */
/* ------------------------------------------------------------------ */
static void hexlin_dndxi(const double *xi, double *dN)
{
    const double u = xi[0], v = xi[1], w = xi[2];
    /* sign pattern of the corner (xi_a, eta_a, zeta_a) for each node */
    static const double sx[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
    static const double sy[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
    static const double sz[8] = {1, 1, 1, 1, -1, -1, -1, -1};
    for (int a = 0; a < 8; a++) {
        dN[3 * a + 0] = sx[a] * 0.125 * (1. + sy[a] * v) * (1. + sz[a] * w);
        dN[3 * a + 1] = sy[a] * 0.125 * (1. + sx[a] * u) * (1. + sz[a] * w);
        dN[3 * a + 2] = sz[a] * 0.125 * (1. + sx[a] * u) * (1. + sy[a] * v);
    }
}

static void hexquad_dndxi(const double *xi, double *dN)
{
    const double u = xi[0], v = xi[1], w = xi[2];
    static const double sx[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
    static const double sy[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
    static const double sz[8] = {1, 1, 1, 1, -1, -1, -1, -1};
    /* corners: N = 1/8 (1+su u)(1+sv v)(1+sw w)(su u + sv v + sw w - 2) */
    for (int a = 0; a < 8; a++) {
        const double pu = 1. + sx[a] * u, pv = 1. + sy[a] * v, pw = 1. + sz[a] * w;
        dN[3 * a + 0] = 0.125 * sx[a] * pv * pw * (2.0 * sx[a] * u + sy[a] * v + sz[a] * w - 1.0);
        dN[3 * a + 1] = 0.125 * sy[a] * pu * pw * (2.0 * sy[a] * v + sx[a] * u + sz[a] * w - 1.0);
        dN[3 * a + 2] = 0.125 * sz[a] * pu * pv * (2.0 * sz[a] * w + sx[a] * u + sy[a] * v - 1.0);
    }
    /* mid-edge nodes 9-12 (zeta=+1), 13-16 (zeta=-1): alternating v-edge / u-edge */
    for (int layer = 0; layer < 2; layer++) {
        const double sw = layer == 0 ? 1.0 : -1.0;
        const double pw = 1.0 + sw * w;
        const int b = 8 + 4 * layer;
        /* node b+0: u=-1 edge along v */
        dN[3 * (b + 0) + 0] = -0.25 * (1.0 - v * v) * pw;
        dN[3 * (b + 0) + 1] = -0.5 * v * (1.0 - u) * pw;
        dN[3 * (b + 0) + 2] = sw * 0.25 * (1.0 - v * v) * (1.0 - u);
        /* node b+1: v=+1 edge along u */
        dN[3 * (b + 1) + 0] = -0.5 * u * (1.0 + v) * pw;
        dN[3 * (b + 1) + 1] = 0.25 * (1.0 - u * u) * pw;
        dN[3 * (b + 1) + 2] = sw * 0.25 * (1.0 - u * u) * (1.0 + v);
        /* node b+2: u=+1 edge along v */
        dN[3 * (b + 2) + 0] = 0.25 * (1.0 - v * v) * pw;
        dN[3 * (b + 2) + 1] = -0.5 * v * (1.0 + u) * pw;
        dN[3 * (b + 2) + 2] = sw * 0.25 * (1.0 - v * v) * (1.0 + u);
        /* node b+3: v=-1 edge along u */
        dN[3 * (b + 3) + 0] = -0.5 * u * (1.0 - v) * pw;
        dN[3 * (b + 3) + 1] = -0.25 * (1.0 - u * u) * pw;
        dN[3 * (b + 3) + 2] = sw * 0.25 * (1.0 - u * u) * (1.0 - v);
    }
    /* vertical mid-edge nodes 17-20 at (u,v) corners (-,-), (-,+), (+,+), (+,-) */
    static const double cu[4] = {-1, -1, 1, 1};
    static const double cv[4] = {-1, 1, 1, -1};
    for (int k = 0; k < 4; k++) {
        const int n = 16 + k;
        const double pu = 1.0 + cu[k] * u, pv = 1.0 + cv[k] * v;
        dN[3 * n + 0] = cu[k] * 0.25 * pv * (1.0 - w * w);
        dN[3 * n + 1] = cv[k] * 0.25 * pu * (1.0 - w * w);
        dN[3 * n + 2] = -0.5 * pu * pv * w;
    }
}

int orc_dndxi(int etype, const double *xi, double *dN)
{
    if (etype == ORC_LSPACE) { hexlin_dndxi(xi, dN); return 8; }
    if (etype == ORC_QSPACE) { hexquad_dndxi(xi, dN); return 20; }
    if (etype == ORC_LTRSPACE) {
        /* N = (l1, l2, l3, 1-l1-l2-l3) fei3dtetlin.C:62-67 */
        static const double t[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, -1, -1, -1};
        memcpy(dN, t, sizeof t);
        return 4;
    }
    return 0;
}

/* FloatMatrix::giveDeterminant floatmatrix.C:1085-1101 (3x3 branch) on J(i,j) */
static double det3_givedet(const double J[9])
{
#define M(r, c) J[3 * (r) + (c)]
    return M(0, 0) * M(1, 1) * M(2, 2) + M(1, 0) * M(2, 1) * M(0, 2) + M(2, 0) * M(0, 1) * M(1, 2)
         - M(0, 2) * M(1, 1) * M(2, 0) - M(1, 2) * M(2, 1) * M(0, 0) - M(2, 2) * M(0, 1) * M(1, 0);
#undef M
}

/* FloatMatrix::beInverseOf floatmatrix.C:768-815 (closed-form 3x3) */
static int inv3_beinverse(const double S[9], double I[9])
{
#define s(r, c) S[3 * ((r) - 1) + ((c) - 1)]
#define o(r, c) I[3 * ((r) - 1) + ((c) - 1)]
    double det = s(1, 1) * s(2, 2) * s(3, 3) + s(1, 2) * s(2, 3) * s(3, 1) +
                 s(1, 3) * s(2, 1) * s(3, 2) - s(1, 3) * s(2, 2) * s(3, 1) -
                 s(2, 3) * s(3, 2) * s(1, 1) - s(3, 3) * s(1, 2) * s(2, 1);
    if (!(fabs(det) > 1.e-30)) return 0;
    o(1, 1) = (s(2, 2) * s(3, 3) - s(2, 3) * s(3, 2)) / det;
    o(2, 1) = (s(2, 3) * s(3, 1) - s(2, 1) * s(3, 3)) / det;
    o(3, 1) = (s(2, 1) * s(3, 2) - s(2, 2) * s(3, 1)) / det;
    o(1, 2) = (s(1, 3) * s(3, 2) - s(1, 2) * s(3, 3)) / det;
    o(2, 2) = (s(1, 1) * s(3, 3) - s(1, 3) * s(3, 1)) / det;
    o(3, 2) = (s(1, 2) * s(3, 1) - s(1, 1) * s(3, 2)) / det;
    o(1, 3) = (s(1, 2) * s(2, 3) - s(1, 3) * s(2, 2)) / det;
    o(2, 3) = (s(1, 3) * s(2, 1) - s(1, 1) * s(2, 3)) / det;
    o(3, 3) = (s(1, 1) * s(2, 2) - s(1, 2) * s(2, 1)) / det;
    return 1;
#undef s
#undef o
}

/* tetrahedron: FEI3dTetLin::evaldNdx fei3dtetlin.C:116-168 (cofactors, abort on detJ<=0) */
static double tet_dndx(const double *X, double *dNdx, int *err)
{
    const double x1 = X[0], y1 = X[1], z1 = X[2], x2 = X[3], y2 = X[4], z2 = X[5];
    const double x3 = X[6], y3 = X[7], z3 = X[8], x4 = X[9], y4 = X[10], z4 = X[11];
    double detJ = ((x4 - x1) * (y2 - y1) * (z3 - z1) - (x4 - x1) * (y3 - y1) * (z2 - z1) +
                   (x3 - x1) * (y4 - y1) * (z2 - z1) - (x2 - x1) * (y4 - y1) * (z3 - z1) +
                   (x2 - x1) * (y3 - y1) * (z4 - z1) - (x3 - x1) * (y2 - y1) * (z4 - z1));
    if (detJ <= 0.0) { if (err) *err = 1; return detJ; }
    double a[12];
    a[0 * 3 + 0] = -((y3 - y2) * (z4 - z2) - (y4 - y2) * (z3 - z2));
    a[1 * 3 + 0] = (y4 - y3) * (z1 - z3) - (y1 - y3) * (z4 - z3);
    a[2 * 3 + 0] = -((y1 - y4) * (z2 - z4) - (y2 - y4) * (z1 - z4));
    a[3 * 3 + 0] = (y2 - y1) * (z3 - z1) - (y3 - y1) * (z2 - z1);
    a[0 * 3 + 1] = -((x4 - x2) * (z3 - z2) - (x3 - x2) * (z4 - z2));
    a[1 * 3 + 1] = (x1 - x3) * (z4 - z3) - (x4 - x3) * (z1 - z3);
    a[2 * 3 + 1] = -((x2 - x4) * (z1 - z4) - (x1 - x4) * (z2 - z4));
    a[3 * 3 + 1] = (x3 - x1) * (z2 - z1) - (x2 - x1) * (z3 - z1);
    a[0 * 3 + 2] = -((x3 - x2) * (y4 - y2) - (x4 - x2) * (y3 - y2));
    a[1 * 3 + 2] = (x4 - x3) * (y1 - y3) - (x1 - x3) * (y4 - y3);
    a[2 * 3 + 2] = -((x1 - x4) * (y2 - y4) - (x2 - x4) * (y1 - y4));
    a[3 * 3 + 2] = (x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1);
    /* answer.times(1./detJ) */
    const double f = 1. / detJ;
    for (int i = 0; i < 12; i++) dNdx[i] = a[i] * f;
    return detJ;
}

/* J(i,j) = sum_k x_k^i dN_k/dxi_j  (jacobianMatrix.beProductOf(coords, dNduvw)) */
static void jacobian(int nen, const double *X, const double *dN, double J[9])
{
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0.;
            for (int k = 0; k < nen; k++) s += X[3 * k + i] * dN[3 * k + j];
            J[3 * i + j] = s;
        }
}

/* FEI3dHexaLin::evaldNdx fei3dhexalin.C:184-200 / FEI3dHexaQuad::evaldNdx fei3dhexaquad.C:284-300:
 * dNdx = dNdxi . inv(J); returns J.giveDeterminant() */
double orc_dndx(int etype, const double *X, const double *xi, double *dNdx, int *err)
{
    if (err) *err = 0;
    if (etype == ORC_LTRSPACE) return tet_dndx(X, dNdx, err);
    int nen = orc_nen(etype);
    double dN[60], J[9], Ji[9];
    orc_dndxi(etype, xi, dN);
    jacobian(nen, X, dN, J);
    if (!inv3_beinverse(J, Ji)) { if (err) *err = 2; return 0.0; }
    for (int k = 0; k < nen; k++)
        for (int j = 0; j < 3; j++) {
            double s = 0.;
            for (int m = 0; m < 3; m++) s += dN[3 * k + m] * Ji[3 * m + j];
            dNdx[3 * k + j] = s;
        }
    return det3_givedet(J);
}

/* FEInterpolation::giveTransformationJacobian feinterpol.C:81-86; tetra fei3dtetlin.C:246-268 */
double orc_detj(int etype, const double *X, const double *xi)
{
    if (etype == ORC_LTRSPACE) {
        double d[12];
        int err = 0;
        double dj = tet_dndx(X, d, &err);
        return dj;
    }
    int nen = orc_nen(etype);
    double dN[60], J[9];
    orc_dndxi(etype, xi, dN);
    jacobian(nen, X, dN, J);
    return det3_givedet(J);
}

/* Structural3DElement::computeBmatrixAt structural3delement.C:64-85: 6 x 3nen */
void orc_bmatrix(int nen, const double *dNdx, double *B)
{
    const int nc = 3 * nen;
    memset(B, 0, sizeof(double) * 6 * nc);
    for (int i = 0; i < nen; i++) {
        const double nx = dNdx[3 * i], ny = dNdx[3 * i + 1], nz = dNdx[3 * i + 2];
        B[0 * nc + 3 * i + 0] = nx;
        B[1 * nc + 3 * i + 1] = ny;
        B[2 * nc + 3 * i + 2] = nz;
        B[4 * nc + 3 * i + 0] = B[3 * nc + 3 * i + 1] = nz;
        B[5 * nc + 3 * i + 0] = B[3 * nc + 3 * i + 2] = ny;
        B[5 * nc + 3 * i + 1] = B[4 * nc + 3 * i + 2] = nx;
    }
}

/* I_dev6 and I6_I6: floatmatrixf.h:954-971 */
static const double IDEV6[36] = {
    2. / 3., -1. / 3., -1. / 3., 0., 0., 0.,
    -1. / 3., 2. / 3., -1. / 3., 0., 0., 0.,
    -1. / 3., -1. / 3., 2. / 3., 0., 0., 0.,
    0., 0., 0., 0.5, 0., 0.,
    0., 0., 0., 0., 0.5, 0.,
    0., 0., 0., 0., 0., 0.5};

/* IsotropicLinearElasticMaterial::initializeFrom/initTangents isolinearelasticmaterial.C:65-88:
 * G = E/(2(1+nu)), K = E/(3(1-2nu)), D = 2G I_dev6 + K I6_I6 */
void orc_isole_D(double E, double nu, double *D)
{
    const double G = E / (2.0 * (1. + nu));
    const double K = E / (3.0 * (1. - 2. * nu));
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) {
            const double ii = (i < 3 && j < 3) ? 1.0 : 0.0;
            D[6 * i + j] = 2 * G * IDEV6[6 * i + j] + K * ii;
        }
}

/* NLStructuralElement::computeStiffnessMatrix nlstructuralelement.C:332-369,446-448:
 * DB = D.B; K += B^T DB dV (plusProductSymmUpper + symmetrized() when the
 * material is symmetric, plusProductUnsym otherwise, floatmatrix.C:691,1243);
 * dV = |detJ| w (structural3delement.C:331-341). */
int orc_element_stiffness(int etype, const double *X, const double *Dgp, int dstride,
                          int symmetric, double *Ke)
{
    const int nen = orc_nen(etype), ngp = orc_ngp(etype), nc = 3 * nen;
    double xi[81], w[27], dNdx[60], B[6 * 60], DB[6 * 60];
    orc_gauss(etype, xi, w);
    memset(Ke, 0, sizeof(double) * nc * nc);
    for (int g = 0; g < ngp; g++) {
        int err = 0;
        orc_dndx(etype, X, xi + 3 * g, dNdx, &err);
        if (err) return -1;
        orc_bmatrix(nen, dNdx, B);
        const double *D = Dgp + (size_t)dstride * g;
        for (int i = 0; i < 6; i++)
            for (int j = 0; j < nc; j++) {
                double s = 0.;
                for (int k = 0; k < 6; k++) s += D[6 * i + k] * B[k * nc + j];
                DB[i * nc + j] = s;
            }
        const double dV = fabs(orc_detj(etype, X, xi + 3 * g)) * w[g];
        for (int i = 0; i < nc; i++)
            for (int j = symmetric ? i : 0; j < nc; j++) {
                double s = 0.;
                for (int k = 0; k < 6; k++) s += B[k * nc + i] * DB[k * nc + j];
                Ke[i * nc + j] += s * dV;
            }
    }
    if (symmetric)
        for (int i = 0; i < nc; i++)
            for (int j = 0; j < i; j++) Ke[i * nc + j] = Ke[j * nc + i];
    return 0;
}

/* ------------------------------------------------------------------ */
/* MisesMat (htype 0, no damage): performPlasticityReturn misesmat.C:181-263
 * (3D branch :221-256) with structuralmaterial.C:1489-1600 helpers;
 * tangent give3dMaterialStiffnessMatrix misesmat.C:496-544.
 * mp = [model, E, nu, sig0, H, ...]                                    */
/* ------------------------------------------------------------------ */
void orc_mises_return(const double *mp, const double *eps, const double *epsp_n, double kappa_n,
                      double *sig, double *epsp_t, double *kappa_t, double *sdev_trial)
{
    const double E = mp[1], nu = mp[2], sig0 = mp[3], H = mp[4];
    const double G = E / (2.0 * (1. + nu));
    const double K = E / (3.0 * (1. - 2. * nu));
    double el[6], pl[6], s[6];
    for (int i = 0; i < 6; i++) { pl[i] = epsp_n[i]; el[i] = eps[i] - pl[i]; }
    double kappa = kappa_n;
    /* computeDeviatoricVolumetricSplit */
    const double vol = el[0] + el[1] + el[2];
    const double mean = vol / 3.0;
    el[0] -= mean; el[1] -= mean; el[2] -= mean;
    /* applyDeviatoricElasticStiffness */
    s[0] = 2. * G * el[0]; s[1] = 2. * G * el[1]; s[2] = 2. * G * el[2];
    s[3] = G * el[3]; s[4] = G * el[4]; s[5] = G * el[5];
    const double pvol = 3 * K * mean;
    for (int i = 0; i < 6; i++) sdev_trial[i] = s[i];
    const double trialS = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2] +
                               2. * s[3] * s[3] + 2. * s[4] * s[4] + 2. * s[5] * s[5]);
    const double yieldValue = sqrt(3. / 2.) * trialS - (sig0 + H * kappa);
    if (yieldValue > 0.) {
        const double dKappa = yieldValue / (H + 3. * G);
        kappa += dKappa;
        /* applyDeviatoricElasticCompliance(s, 0.5) */
        double dpl[6];
        dpl[0] = 1. / (2. * 0.5) * s[0]; dpl[1] = 1. / (2. * 0.5) * s[1]; dpl[2] = 1. / (2. * 0.5) * s[2];
        dpl[3] = 1. / 0.5 * s[3]; dpl[4] = 1. / 0.5 * s[4]; dpl[5] = 1. / 0.5 * s[5];
        const double f = sqrt(3. / 2.) * dKappa / trialS;
        for (int i = 0; i < 6; i++) pl[i] += f * dpl[i];
        const double sc = 1. - sqrt(6.) * G * dKappa / trialS;
        for (int i = 0; i < 6; i++) s[i] *= sc;
    }
    /* computeDeviatoricVolumetricSum; damage factor (1-omega) = 1 */
    s[0] += pvol; s[1] += pvol; s[2] += pvol;
    for (int i = 0; i < 6; i++) { sig[i] = s[i] * (1 - 0.0); epsp_t[i] = pl[i]; }
    *kappa_t = kappa;
}

void orc_mises_tangent(const double *mp, double kappa_n, double kappa_t,
                       const double *sdev_trial, double *D)
{
    const double E = mp[1], nu = mp[2], sig0 = mp[3], H = mp[4];
    const double G = E / (2.0 * (1. + nu));
    orc_isole_D(E, nu, D);
    const double dKappa = kappa_t - kappa_n;
    if (dKappa <= 0.0) return;
    const double sigmaY = sig0 + H * kappa_n;
    const double *s = sdev_trial;
    const double trialS = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2] +
                               2. * s[3] * s[3] + 2. * s[4] * s[4] + 2. * s[5] * s[5]);
    const double factor = -2. * sqrt(6.) * G * G / trialS;
    const double factor1 = factor * sigmaY / ((H + 3. * G) * trialS * trialS);
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) D[6 * i + j] += factor1 * (s[i] * s[j]);
    const double factor2 = factor * dKappa;
    for (int i = 0; i < 36; i++) D[i] += factor2 * IDEV6[i];
}

/* ------------------------------------------------------------------ */
/* mesh helpers                                                        */
/* ------------------------------------------------------------------ */
static void gather_coords(const orc_mesh *m, int e, double *X)
{
    const int nen = orc_nen(m->etype);
    for (int a = 0; a < nen; a++) {
        const int n = m->conn[(size_t)e * nen + a];
        X[3 * a + 0] = m->coords[3 * (size_t)n + 0];
        X[3 * a + 1] = m->coords[3 * (size_t)n + 1];
        X[3 * a + 2] = m->coords[3 * (size_t)n + 2];
    }
}

/* Element::giveLocationArray element.C:429-455 with the u,v,w dof mask
 * (structural3delement.C:300): free dofs give their equation, others 0 */
static void location(const orc_mesh *m, int e, int *loc, int *ploc)
{
    const int nen = orc_nen(m->etype);
    for (int a = 0; a < nen; a++) {
        const int n = m->conn[(size_t)e * nen + a];
        for (int d = 0; d < 3; d++) {
            const int q = m->eqn[3 * (size_t)n + d];
            loc[3 * a + d] = q > 0 ? q : 0;
            if (ploc) ploc[3 * a + d] = q < 0 ? -q : 0;
        }
    }
}

static int cmp_int(const void *a, const void *b)
{
    const int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

/* Node-to-element adjacency (elements in ascending order) */
static void node_elements(const orc_mesh *m, int **ptr_out, int **lst_out)
{
    const int nen = orc_nen(m->etype);
    int *ptr = (int *)calloc((size_t)m->nnode + 1, sizeof(int));
    for (size_t i = 0; i < (size_t)m->nelem * nen; i++) ptr[m->conn[i] + 1]++;
    for (int n = 0; n < m->nnode; n++) ptr[n + 1] += ptr[n];
    int *fill = (int *)malloc(sizeof(int) * (size_t)m->nnode);
    memcpy(fill, ptr, sizeof(int) * (size_t)m->nnode);
    int *lst = (int *)malloc(sizeof(int) * (size_t)ptr[m->nnode] + 1);
    for (int e = 0; e < m->nelem; e++)
        for (int a = 0; a < nen; a++) lst[fill[m->conn[(size_t)e * nen + a]]++] = e;
    free(fill);
    *ptr_out = ptr;
    *lst_out = lst;
}

/* CompCol::buildInternalStructure compcol.C:174-272: the pattern is the union
 * of loc x loc over all elements, free equations only, sorted columns.
 * Restated row-wise (full, unsymmetric storage).  Each node's free rows
 * share one column list: the free equations of all nodes it shares an
 * element with. */
static long long build_rows(const orc_mesh *m, long long *rowptr, int *colind)
{
    const int nen = orc_nen(m->etype);
    int *nptr, *nlst;
    node_elements(m, &nptr, &nlst);
    int *mark = (int *)malloc(sizeof(int) * (size_t)m->nnode);
    for (int n = 0; n < m->nnode; n++) mark[n] = -1;
    int *cols = (int *)malloc(sizeof(int) * 4096);
    long long total = 0;
    if (rowptr) rowptr[0] = 0;
    /* first pass: lengths, indexed by equation */
    long long *len = rowptr ? (long long *)calloc((size_t)m->neq + 1, sizeof(long long)) : NULL;
    for (int a = 0; a < m->nnode; a++) {
        int nfree_a = 0;
        for (int d = 0; d < 3; d++) nfree_a += m->eqn[3 * (size_t)a + d] > 0;
        if (!nfree_a) continue;
        int nc = 0;
        for (int k = nptr[a]; k < nptr[a + 1]; k++) {
            const int e = nlst[k];
            for (int b = 0; b < nen; b++) {
                const int nb = m->conn[(size_t)e * nen + b];
                if (mark[nb] == a) continue;
                mark[nb] = a;
                for (int d = 0; d < 3; d++) {
                    const int q = m->eqn[3 * (size_t)nb + d];
                    if (q > 0) cols[nc++] = q - 1;
                }
            }
        }
        total += (long long)nc * nfree_a;
        if (rowptr) {
            qsort(cols, (size_t)nc, sizeof(int), cmp_int);
            for (int d = 0; d < 3; d++) {
                const int q = m->eqn[3 * (size_t)a + d];
                if (q > 0) len[q] = nc;
            }
        }
    }
    if (rowptr) {
        for (int r = 0; r < m->neq; r++) rowptr[r + 1] = rowptr[r] + len[r + 1];
        for (int n = 0; n < m->nnode; n++) mark[n] = -1;
        for (int a = 0; a < m->nnode; a++) {
            int nc = 0;
            for (int k = nptr[a]; k < nptr[a + 1]; k++) {
                const int e = nlst[k];
                for (int b = 0; b < nen; b++) {
                    const int nb = m->conn[(size_t)e * nen + b];
                    if (mark[nb] == a) continue;
                    mark[nb] = a;
                    for (int d = 0; d < 3; d++) {
                        const int q = m->eqn[3 * (size_t)nb + d];
                        if (q > 0) cols[nc++] = q - 1;
                    }
                }
            }
            qsort(cols, (size_t)nc, sizeof(int), cmp_int);
            for (int d = 0; d < 3; d++) {
                const int q = m->eqn[3 * (size_t)a + d];
                if (q > 0) memcpy(colind + rowptr[q - 1], cols, sizeof(int) * (size_t)nc);
            }
        }
        free(len);
    }
    free(cols);
    free(mark);
    free(nptr);
    free(nlst);
    return total;
}

long long orc_csr_count(const orc_mesh *m) { return build_rows(m, NULL, NULL); }

int orc_csr_fill(const orc_mesh *m, long long *rowptr, int *colind)
{
    build_rows(m, rowptr, colind);
    return 0;
}

static long long find_pos(const long long *rowptr, const int *colind, int r, int c)
{
    long long lo = rowptr[r], hi = rowptr[r + 1] - 1;
    while (lo <= hi) {
        const long long mid = (lo + hi) >> 1;
        if (colind[mid] == c) return mid;
        if (colind[mid] < c) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

/* Constitutive matrix at each Gauss point of element e.  IsoLE: D (symmetric
 * material, material.h isCharacteristicMtrxSymmetric true); MisesMat: the
 * consistent tangent from the temp state (symmetric flag false, misesmat.h:125). */
static int element_D(const orc_mesh *m, int e, const double *state, int tangent, double *Dgp, int *sym)
{
    const double *mp = m->mat + (size_t)ORC_MATP * m->emat[e];
    const int ngp = orc_ngp(m->etype);
    if ((int)mp[0] == ORC_MAT_ISOLE) {
        for (int g = 0; g < ngp; g++) orc_isole_D(mp[1], mp[2], Dgp + 36 * g);
        *sym = 1;
        return 0;
    }
    *sym = 0;
    for (int g = 0; g < ngp; g++) {
        const double *st = state + (size_t)ORC_GPSTATE * ((size_t)e * ngp + g);
        if (tangent && state)
            orc_mises_tangent(mp, st[6], st[13], st + 14, Dgp + 36 * g);
        else
            orc_isole_D(mp[1], mp[2], Dgp + 36 * g);
    }
    return 0;
}

/* EngngModel::assemble engngm.C:889-1037 with SymCompCol/CompCol::assemble
 * (symcompcol.C:232, compcol.C:275): elements in ascending order, each K_e
 * entry added at (loc_i, loc_j) for loc > 0. */
int orc_assemble_stiffness(const orc_mesh *m, const long long *rowptr, const int *colind,
                           double *val, const double *state, int tangent)
{
    const int nen = orc_nen(m->etype), nc = 3 * nen, ngp = orc_ngp(m->etype);
    double X[60], *Ke = (double *)malloc(sizeof(double) * nc * nc), *Dgp = (double *)malloc(sizeof(double) * 36 * ngp);
    int loc[60];
    memset(val, 0, sizeof(double) * (size_t)rowptr[m->neq]);
    int rc = 0;
    for (int e = 0; e < m->nelem && !rc; e++) {
        int sym = 1;
        gather_coords(m, e, X);
        element_D(m, e, state, tangent, Dgp, &sym);
        if (orc_element_stiffness(m->etype, X, Dgp, 36, sym, Ke)) { rc = -1 - e; break; }
        location(m, e, loc, NULL);
        for (int i = 0; i < nc; i++) {
            if (!loc[i]) continue;
            for (int j = 0; j < nc; j++) {
                if (!loc[j]) continue;
                const long long p = find_pos(rowptr, colind, loc[i] - 1, loc[j] - 1);
                if (p < 0) { rc = -2; break; }
                val[p] += Ke[i * nc + j];
            }
        }
    }
    free(Ke);
    free(Dgp);
    return rc;
}

int orc_assemble_stiffness_omp(const orc_mesh *m, const long long *rowptr, const int *colind, double *val,
                               const double *state, int tangent)
{
    const int nen = orc_nen(m->etype), nc = 3 * nen, ngp = orc_ngp(m->etype);
    memset(val, 0, sizeof(double) * (size_t)rowptr[m->neq]);
    int rc = 0;
#pragma omp parallel
    {
        double X[60], *Ke = (double *)malloc(sizeof(double) * nc * nc), *Dgp = (double *)malloc(sizeof(double) * 36 * ngp);
        int loc[60];
#pragma omp for schedule(static)
        for (int e = 0; e < m->nelem; e++) {
            int sym = 1;
            gather_coords(m, e, X);
            element_D(m, e, state, tangent, Dgp, &sym);
            const int bad = orc_element_stiffness(m->etype, X, Dgp, 36, sym, Ke);
            location(m, e, loc, NULL);
#pragma omp critical
            {
                if (bad) rc = -1 - e;
                for (int i = 0; i < nc && !bad; i++) {
                    if (!loc[i]) continue;
                    for (int j = 0; j < nc; j++) {
                        if (!loc[j]) continue;
                        const long long p = find_pos(rowptr, colind, loc[i] - 1, loc[j] - 1);
                        if (p < 0) { rc = -2; break; }
                        val[p] += Ke[i * nc + j];
                    }
                }
            }
        }
        free(Ke);
        free(Dgp);
    }
    return rc;
}

/* CompCol::times compcol.C:128 restated row-wise: y = A x */
void orc_spmv(int n, const long long *rowptr, const int *colind, const double *val,
              const double *x, double *y)
{
    for (int r = 0; r < n; r++) {
        double s = 0.;
        for (long long k = rowptr[r]; k < rowptr[r + 1]; k++) s += val[k] * x[colind[k]];
        y[r] = s;
    }
}

/* SymCompCol::times symcompcol.C:201-221 restated on the full symmetric CSR:
 * column j of the lower-triangle storage is row j's entries with col >= j.
 * y[i] = (sum over j<i, ascending j, of A(i,j) x[j]) + (sum over k>i,
 * ascending, of A(k,i) x[k] + A(i,i) x[i]).                             */
void orc_spmv_symcompcol(int n, const long long *rowptr, const int *colind, const double *val,
                         const double *x, double *y)
{
    for (int r = 0; r < n; r++) y[r] = 0.0;
    for (int j = 0; j < n; j++) {
        const double rhs = x[j];
        double sum = 0.0, dg = 0.0;
        for (long long t = rowptr[j]; t < rowptr[j + 1]; t++) {
            const int i = colind[t];
            if (i < j) continue;
            if (i == j) { dg = val[t]; continue; }
            y[i] += val[t] * rhs;
            sum += val[t] * x[i];
        }
        y[j] += sum + dg * rhs;
    }
}

static double dotp(int n, const double *a, const double *b)
{
    double s = 0.;
    for (int i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

/* IMLSolver::solve imlsolver.C:102-146 -> CG iml/cg.h:22-72 with
 * DiagPreconditioner diagpre.C:41-68 (precond 1) or none (precond 0).
 * Return 0 on convergence, 1 when maxit is reached, <0 on error. */
int orc_pcg(int n, const long long *rowptr, const int *colind, const double *val,
            const double *b, double *x, double tol, int maxit, int precond,
            int *iters, double *resid)
{
    return orc_pcg_hist(n, rowptr, colind, val, b, x, tol, maxit, precond, iters, resid, NULL, 0);
}

/* orc_pcg that also stores ||r_k|| / ||b|| for k = 0, 1, ... (the values
 * cg.h tests against tol) in hist[0..cap) */
int orc_pcg_hist(int n, const long long *rowptr, const int *colind, const double *val,
                 const double *b, double *x, double tol, int maxit, int precond,
                 int *iters, double *resid, double *hist, int cap)
{
    /* precond bit 0: DiagPreconditioner; bit 4: SpMV in SymCompCol order */
    void (*mv)(int, const long long *, const int *, const double *, const double *, double *) =
        (precond & 16) ? orc_spmv_symcompcol : orc_spmv;
    precond &= 15;
    double *dinv = (double *)malloc(sizeof(double) * (size_t)n);
    double *r = (double *)malloc(sizeof(double) * (size_t)n), *z = (double *)malloc(sizeof(double) * (size_t)n);
    double *p = (double *)malloc(sizeof(double) * (size_t)n), *q = (double *)malloc(sizeof(double) * (size_t)n);
    int rc = 1;
    for (int i = 0; i < n; i++) {
        double d = 1.0;
        if (precond == 1) {
            const long long k = find_pos(rowptr, colind, i, i);
            d = k >= 0 ? val[k] : 0.0;
            if (d == 0) { rc = -1; goto done; }
        }
        dinv[i] = 1. / d;
    }
    {
        double normb = sqrt(dotp(n, b, b));
        mv(n, rowptr, colind, val, x, q);
        for (int i = 0; i < n; i++) r[i] = b[i] - q[i];
        if (normb == 0.0) normb = 1;
        double res = sqrt(dotp(n, r, r)) / normb, rho_1 = 0.0;
        if (hist && cap > 0) hist[0] = res;
        if (res <= tol) { *resid = res; *iters = 0; rc = 0; goto done; }
        for (int it = 1; it <= maxit; it++) {
            for (int i = 0; i < n; i++) z[i] = r[i] * dinv[i];
            const double rho = dotp(n, r, z);
            if (it == 1) memcpy(p, z, sizeof(double) * (size_t)n);
            else {
                const double beta = rho / rho_1;
                for (int i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
            }
            mv(n, rowptr, colind, val, p, q);
            const double alpha = rho / dotp(n, p, q);
            for (int i = 0; i < n; i++) x[i] += alpha * p[i];
            for (int i = 0; i < n; i++) r[i] -= alpha * q[i];
            res = sqrt(dotp(n, r, r)) / normb;
            if (hist && it < cap) hist[it] = res;
            *resid = res;
            *iters = it;
            if (res <= tol) { rc = 0; goto done; }
            rho_1 = rho;
        }
        rc = 1;
    }
done:
    free(dinv); free(r); free(z); free(p); free(q);
    return rc;
}

/* ------------------------------------------------------------------ */
/* internal forces: NLStructuralElement::giveInternalForcesVector
 * nlstructuralelement.C:151-241 (nlGeometry 0): u_e = computeVectorOf(VM_Total)
 * (element.C:103), eps = B u_e, sigma = material(eps), f_e += B^T sigma dV.
 * Assembly and EBE norms: EngngModel::assembleVectorFromElements
 * engngm.C:1351-1420 (FloatArray::assembleSquared floatarray.C:637 over the
 * element's u,v,w dof ids).                                             */
/* ------------------------------------------------------------------ */
void orc_state_init(const orc_mesh *m, double *state)
{
    memset(state, 0, sizeof(double) * ORC_GPSTATE * (size_t)m->nelem * orc_ngp(m->etype));
}

/* MisesMatStatus::updateYourself misesmat.C:720-728 (+ StructuralMaterialStatus stress) */
void orc_state_commit(const orc_mesh *m, double *state)
{
    const size_t n = (size_t)m->nelem * orc_ngp(m->etype);
    for (size_t g = 0; g < n; g++) {
        double *st = state + ORC_GPSTATE * g;
        for (int i = 0; i < 6; i++) st[i] = st[7 + i];
        st[6] = st[13];
        for (int i = 0; i < 6; i++) st[32 + i] = st[20 + i];
    }
}

/* MisesMatStatus::initTempStatus misesmat.C:707-715 (+ structuralms.C:121-141)
 * for every Gauss point: temp plastic strain, kappa and stress := committed,
 * trial deviator cleared (the temp strain has no committed copy; the next
 * stress evaluation rewrites it) */
void orc_state_restart(const orc_mesh *m, double *state)
{
    const size_t n = (size_t)m->nelem * orc_ngp(m->etype);
    for (size_t g = 0; g < n; g++) {
        double *st = state + ORC_GPSTATE * g;
        for (int i = 0; i < 6; i++) {
            st[7 + i] = st[i];
            st[14 + i] = 0.0;
            st[20 + i] = st[32 + i];
        }
        st[13] = st[6];
    }
}

/* one element of orc_internal_forces: f_e = sum_gp B^T sigma dV with the
 * Gauss-point state update; loc / ploc its free / prescribed equations */
static int elem_internal_forces(const orc_mesh *m, int e, const double *u_free, const double *u_presc,
                                double *state, int update_state, const double *xi, const double *w, double *fe,
                                int *loc, int *ploc)
{
    const int nen = orc_nen(m->etype), nc = 3 * nen, ngp = orc_ngp(m->etype);
    double X[60], dNdx[60], B[6 * 60], ue[60];
    const double *mp = m->mat + (size_t)ORC_MATP * m->emat[e];
    gather_coords(m, e, X);
    location(m, e, loc, ploc);
    for (int i = 0; i < nc; i++)
        ue[i] = loc[i] ? u_free[loc[i] - 1] : (ploc[i] && u_presc ? u_presc[ploc[i] - 1] : 0.0);
    memset(fe, 0, sizeof(double) * nc);
    for (int g = 0; g < ngp; g++) {
        int err = 0;
        orc_dndx(m->etype, X, xi + 3 * g, dNdx, &err);
        if (err) return -1 - e;
        orc_bmatrix(nen, dNdx, B);
        double eps[6], sig[6];
        for (int i = 0; i < 6; i++) {
            double s = 0.;
            for (int j = 0; j < nc; j++) s += B[i * nc + j] * ue[j];
            eps[i] = s;
        }
        double *st = state ? state + (size_t)ORC_GPSTATE * ((size_t)e * ngp + g) : NULL;
        if ((int)mp[0] == ORC_MAT_MISES) {
            double epsp_t[6], kappa_t, sdev[6];
            orc_mises_return(mp, eps, st ? st : (const double[6]){0, 0, 0, 0, 0, 0},
                             st ? st[6] : 0.0, sig, epsp_t, &kappa_t, sdev);
            if (st && update_state) {
                for (int i = 0; i < 6; i++) { st[7 + i] = epsp_t[i]; st[14 + i] = sdev[i]; }
                st[13] = kappa_t;
            }
        } else {
            /* LinearElasticMaterial::giveRealStressVector_3d linearelasticmaterial.C:102-124 */
            double D[36];
            orc_isole_D(mp[1], mp[2], D);
            for (int i = 0; i < 6; i++) {
                double s = 0.;
                for (int k = 0; k < 6; k++) s += D[6 * i + k] * eps[k];
                sig[i] = s;
            }
        }
        if (st && update_state)
            for (int i = 0; i < 6; i++) { st[20 + i] = sig[i]; st[26 + i] = eps[i]; }
        const double dV = fabs(orc_detj(m->etype, X, xi + 3 * g)) * w[g];
        for (int i = 0; i < nc; i++) {
            double s = 0.;
            for (int k = 0; k < 6; k++) s += B[k * nc + i] * sig[k];
            fe[i] += s * dV;
        }
    }
    return 0;
}

int orc_internal_forces(const orc_mesh *m, const double *u_free, const double *u_presc,
                        double *state, int update_state,
                        double *f_free, double *f_presc, double *ebe_norm3)
{
    const int nc = 3 * orc_nen(m->etype);
    double xi[81], w[27], fe[60];
    int loc[60], ploc[60];
    orc_gauss(m->etype, xi, w);
    if (f_free) memset(f_free, 0, sizeof(double) * (size_t)m->neq);
    if (f_presc) memset(f_presc, 0, sizeof(double) * (size_t)m->npeq);
    if (ebe_norm3) ebe_norm3[0] = ebe_norm3[1] = ebe_norm3[2] = 0.0;
    for (int e = 0; e < m->nelem; e++) {
        const int rc = elem_internal_forces(m, e, u_free, u_presc, state, update_state, xi, w, fe, loc, ploc);
        if (rc) return rc;
        for (int i = 0; i < nc; i++) {
            if (loc[i] && f_free) f_free[loc[i] - 1] += fe[i];
            if (ploc[i] && f_presc) f_presc[ploc[i] - 1] += fe[i];
            if (ebe_norm3) ebe_norm3[i % 3] += fe[i] * fe[i];
        }
    }
    return 0;
}

/* The reference's OpenMP build (USE_OPENMP_PARALLEL) of the two element
 * loops: EngngModel::assembleVectorFromElements engngm.C:1375-1408 and
 * EngngModel::assemble engngm.C:902-923 -- the elements shared out by
 * "omp parallel for", each element's vector or matrix scattered inside one
 * "omp critical" section.  The order of the additions then depends on the
 * schedule, as it does there: these are timing baselines of the CPU path,
 * not the restatement the parity tests compare with. */
int orc_omp_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int orc_internal_forces_omp(const orc_mesh *m, const double *u_free, const double *u_presc, double *state,
                            int update_state, double *f_free, double *f_presc, double *ebe_norm3)
{
    const int nc = 3 * orc_nen(m->etype);
    double xi[81], w[27];
    orc_gauss(m->etype, xi, w);
    if (f_free) memset(f_free, 0, sizeof(double) * (size_t)m->neq);
    if (f_presc) memset(f_presc, 0, sizeof(double) * (size_t)m->npeq);
    if (ebe_norm3) ebe_norm3[0] = ebe_norm3[1] = ebe_norm3[2] = 0.0;
    int rc = 0;
#pragma omp parallel for schedule(static)
    for (int e = 0; e < m->nelem; e++) {
        double fe[60];
        int loc[60], ploc[60];
        const int r = elem_internal_forces(m, e, u_free, u_presc, state, update_state, xi, w, fe, loc, ploc);
#pragma omp critical
        {
            if (r) rc = r;
            for (int i = 0; i < nc; i++) {
                if (loc[i] && f_free) f_free[loc[i] - 1] += fe[i];
                if (ploc[i] && f_presc) f_presc[ploc[i] - 1] += fe[i];
                if (ebe_norm3) ebe_norm3[i % 3] += fe[i] * fe[i];
            }
        }
    }
    return rc;
}

/* ------------------------------------------------------------------ */
/* surface traction: StructuralElement::computeBoundarySurfaceLoadVector
 * structuralelement.C:110-164; hexa surfaces fei3dhexalin.C:447-458,532-566,
 * 594-613 with a 2x2 Gauss rule (feinterpol.C:123, order 0+1 on _Square ->
 * 4 points, gaussintegrationrule.C:418); tetra surfaces fei3dtetlin.C:361-368,
 * 414-459 with the 1-point triangle rule.                              */
/* ------------------------------------------------------------------ */
static const int HEX_SURF[6][4] = {{1, 4, 3, 2}, {5, 6, 7, 8}, {1, 2, 6, 5},
                                   {2, 3, 7, 6}, {3, 4, 8, 7}, {4, 1, 5, 8}};
static const int TET_SURF[4][3] = {{1, 3, 2}, {1, 2, 4}, {2, 3, 4}, {1, 4, 3}};

/* shape functions: FEI3dHexaLin::evalN fei3dhexalin.C:66-80, FEI3dTetLin::evalN
 * fei3dtetlin.C:62-67, FEI3dHexaQuad::evalN fei3dhexaquad.C:128-189 (node
 * order as in orc_dndxi) */
static void orc_n(int etype, const double *xi, double *N)
{
    const double u = xi[0], v = xi[1], w = xi[2];
    static const double sx[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
    static const double sy[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
    static const double sz[8] = {1, 1, 1, 1, -1, -1, -1, -1};
    if (etype == ORC_LTRSPACE) {
        N[0] = u; N[1] = v; N[2] = w; N[3] = 1.0 - u - v - w;
        return;
    }
    if (etype == ORC_LSPACE) {
        for (int a = 0; a < 8; a++) N[a] = 0.125 * (1. + sx[a] * u) * (1. + sy[a] * v) * (1. + sz[a] * w);
        return;
    }
    for (int a = 0; a < 8; a++)
        N[a] = 0.125 * (1. + sx[a] * u) * (1. + sy[a] * v) * (1. + sz[a] * w) * (sx[a] * u + sy[a] * v + sz[a] * w - 2.);
    for (int layer = 0; layer < 2; layer++) {
        const double pw = 1.0 + (layer == 0 ? 1.0 : -1.0) * w;
        const int b = 8 + 4 * layer;
        N[b + 0] = 0.25 * (1.0 - v * v) * (1.0 - u) * pw;
        N[b + 1] = 0.25 * (1.0 - u * u) * (1.0 + v) * pw;
        N[b + 2] = 0.25 * (1.0 - v * v) * (1.0 + u) * pw;
        N[b + 3] = 0.25 * (1.0 - u * u) * (1.0 - v) * pw;
    }
    static const double cu[4] = {-1, -1, 1, 1};
    static const double cv[4] = {-1, 1, 1, -1};
    for (int k = 0; k < 4; k++) N[16 + k] = 0.25 * (1.0 + cu[k] * u) * (1.0 + cv[k] * v) * (1.0 - w * w);
}

/* StructuralElement::computeBodyLoadVectorAt structuralelement.C:262-296
 * (DeadWeight): f_e = sum over the default Gauss rule of N^T b dV rho, with
 * rho the element's density and b the global force per unit mass */
int orc_body_load(const orc_mesh *m, int nel, const int *elem, const double *dens, const double *b,
                  double *f_free, double *f_presc)
{
    const int nen = orc_nen(m->etype);
    double xi[81], w[27], X[60], N[20];
    const int ngp = orc_gauss(m->etype, xi, w);
    for (int k = 0; k < nel; k++) {
        const int e = elem[k];
        gather_coords(m, e, X);
        double fe[60] = {0};
        for (int g = 0; g < ngp; g++) {
            orc_n(m->etype, xi + 3 * g, N);
            const double dV = fabs(orc_detj(m->etype, X, xi + 3 * g)) * w[g];
            for (int a = 0; a < nen; a++)
                for (int d = 0; d < 3; d++) fe[3 * a + d] += N[a] * b[d] * dV * dens[k];
        }
        for (int a = 0; a < nen; a++)
            for (int d = 0; d < 3; d++) {
                const int q = m->eqn[3 * (size_t)m->conn[(size_t)e * nen + a] + d];
                if (q > 0 && f_free) f_free[q - 1] += fe[3 * a + d];
                if (q < 0 && f_presc) f_presc[-q - 1] += fe[3 * a + d];
            }
    }
    return 0;
}

int orc_surface_load(const orc_mesh *m, int nface, const int *elem, const int *surf,
                     const double *traction, double *f_free, double *f_presc)
{
    const int nen = orc_nen(m->etype);
    double X[60];
    for (int f = 0; f < nface; f++) {
        const int e = elem[f], s = surf[f];
        const double *t = traction + 3 * f;
        gather_coords(m, e, X);
        int sn[4], nsn;
        double fe[12] = {0};
        if (m->etype == ORC_LSPACE) {
            if (s < 1 || s > 6) return -1;
            nsn = 4;
            for (int k = 0; k < 4; k++) sn[k] = HEX_SURF[s - 1][k] - 1;
            for (int i = 0; i < 2; i++)
                for (int j = 0; j < 2; j++) {
                    const double ksi = GL2_X[i], eta = GL2_X[j], wgt = GL2_W[i] * GL2_W[j];
                    const double N[4] = {(1. + ksi) * (1. + eta) * 0.25, (1. - ksi) * (1. + eta) * 0.25,
                                         (1. - ksi) * (1. - eta) * 0.25, (1. + ksi) * (1. - eta) * 0.25};
                    const double dk[4] = {(1. + eta), -(1. + eta), -(1. - eta), (1. - eta)};
                    const double de[4] = {(1. + ksi), (1. - ksi), -(1. - ksi), -(1. + ksi)};
                    double a[3] = {0, 0, 0}, b[3] = {0, 0, 0};
                    for (int k = 0; k < 4; k++)
                        for (int d = 0; d < 3; d++) {
                            a[d] += dk[k] * X[3 * sn[k] + d];
                            b[d] += de[k] * X[3 * sn[k] + d];
                        }
                    const double c0 = a[1] * b[2] - a[2] * b[1], c1 = a[2] * b[0] - a[0] * b[2], c2 = a[0] * b[1] - a[1] * b[0];
                    const double dA = sqrt(c0 * c0 + c1 * c1 + c2 * c2) * 0.0625 * wgt;
                    for (int k = 0; k < 4; k++)
                        for (int d = 0; d < 3; d++) fe[3 * k + d] += N[k] * t[d] * dA;
                }
        } else if (m->etype == ORC_LTRSPACE) {
            if (s < 1 || s > 4) return -1;
            nsn = 3;
            for (int k = 0; k < 3; k++) sn[k] = TET_SURF[s - 1][k] - 1;
            const double l1 = 1. / 3., l2 = 1. / 3., wgt = 0.5;
            const double N[3] = {l1, l2, 1. - l1 - l2};
            double a[3], b[3];
            for (int d = 0; d < 3; d++) {
                a[d] = X[3 * sn[1] + d] - X[3 * sn[0] + d];
                b[d] = X[3 * sn[2] + d] - X[3 * sn[0] + d];
            }
            const double c0 = a[1] * b[2] - a[2] * b[1], c1 = a[2] * b[0] - a[0] * b[2], c2 = a[0] * b[1] - a[1] * b[0];
            const double dA = sqrt(c0 * c0 + c1 * c1 + c2 * c2) * wgt;
            for (int k = 0; k < 3; k++)
                for (int d = 0; d < 3; d++) fe[3 * k + d] += N[k] * t[d] * dA;
        } else {
            return -2;
        }
        for (int k = 0; k < nsn; k++) {
            const int n = m->conn[(size_t)e * nen + sn[k]];
            for (int d = 0; d < 3; d++) {
                const int q = m->eqn[3 * (size_t)n + d];
                if (q > 0 && f_free) f_free[q - 1] += fe[3 * k + d];
                if (q < 0 && f_presc) f_presc[-q - 1] += fe[3 * k + d];
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* EngngModel::assembleExtrapolatedForces engngm.C:1536-1610 with the
 * TangentStiffnessMatrix of the current Gauss-point state: per element
 * K_e * du_e, where du_e holds the prescribed increments (free entries are
 * zero at step start), assembled into the free rows: f = sum_e K_e du_e.
 * The caller negates it (staticstructural.C:267).                      */
int orc_extrapolated_forces(const orc_mesh *m, const double *state, const double *du_presc, double *f)
{
    const int nen = orc_nen(m->etype), nc = 3 * nen, ngp = orc_ngp(m->etype);
    double X[60], du[60], *Ke = (double *)malloc(sizeof(double) * nc * nc), *Dgp = (double *)malloc(sizeof(double) * 36 * ngp);
    int loc[60], ploc[60], rc = 0;
    memset(f, 0, sizeof(double) * (size_t)m->neq);
    for (int e = 0; e < m->nelem; e++) {
        location(m, e, loc, ploc);
        int any = 0;
        for (int i = 0; i < nc; i++) {
            du[i] = ploc[i] ? du_presc[ploc[i] - 1] : 0.0;
            if (du[i] != 0.0) any = 1;
        }
        if (!any) continue;  /* K_e * 0 contributes exact zeros */
        int sym = 1;
        gather_coords(m, e, X);
        element_D(m, e, state, 1, Dgp, &sym);
        if (orc_element_stiffness(m->etype, X, Dgp, 36, sym, Ke)) { rc = -1 - e; break; }
        for (int i = 0; i < nc; i++) {   /* FloatArray::beProductOf */
            if (!loc[i]) continue;
            double s = 0.0;
            for (int j = 0; j < nc; j++) s += Ke[i * nc + j] * du[j];
            f[loc[i] - 1] += s;
        }
    }
    free(Ke);
    free(Dgp);
    return rc;
}

/* StaticStructural::solveYourselfAt staticstructural.C:224-358 with
 * NRSolver::solve nrsolver.C:205-404 and the per-dof-id force criterion of
 * checkConvergence nrsolver.C:594-885 (rtold off):
 *   initial guess (initial_guess = 1, IG_Tangent, the default at
 *   staticstructural.C:130,247-297): f = -sum_e K_e du_e over the prescribed
 *   increments du = u_presc - u_presc_prev; if |f| > 0: K = tangent,
 *   K inc = f (CG from zero), X += inc;
 *   K = tangent of the state at step start (NonLinearLhs before the first
 *   residual); for nite = 0..:
 *     F = f_int(X, prescribed); rhs = RT - F;
 *     err_d = sqrt(sum_{eq of dof d} rhs^2 / (sum RT^2 + eNorm_d)) (absolute
 *       when the denominator < 1e-6); converged when all err_d <= rtolf and
 *       not (nite == 0 and err > 0);
 *     if nite > 0 and manrmsteps > 0 and nite % manrmsteps == 0: K = tangent;
 *     K ddX = rhs (IML CG + Jacobi from the previous ddX, tolerance lstol);
 *     X += ddX.
 * X (neq) is updated in place; state holds the Gauss-point state (temp
 * fields updated, committed fields untouched).  out[0] Newton iterations,
 * out[1] CG iterations (initial guess included), out[2] status (1
 * converged, -1 maxiter, -2 out of range).  Returns the final max force
 * error in *err.  u_presc_prev may be NULL (all zero, the first step). */
/* LineSearchNM::search linesearch.C:163-245: next trial step length from
 * the ratios prod[1..istep] at eta[1..istep] (1-based, as the reference) */
static void orc_ls_search(int istep, double *prod, double *eta, double amp, double maxeta, double mineta, int *ico)
{
    int ineg = 0;
    double etaneg = 1.0, etamax = 0.0;
    for (int i = 1; i <= istep; i++) {
        if (eta[i] > etamax) etamax = eta[i];
        if (prod[i] >= 0.0) continue;
        if (eta[i] >= etaneg) continue;
        etaneg = eta[i];
        ineg = i;
    }
    if (ineg) {
        int ipos = 1;
        for (int i = 1; i <= istep; i++) {
            if (prod[i] <= 0.0) continue;
            if (eta[i] > eta[ineg]) continue;
            if (eta[i] < eta[ipos]) continue;
            ipos = i;
        }
        double etaint = (prod[ineg] * eta[ipos] - prod[ipos] * eta[ineg]) / (prod[ineg] - prod[ipos]);
        const double etaalt = eta[ipos] + 0.2 * (eta[ineg] - eta[ipos]);
        if (etaalt > etaint) etaint = etaalt;
        if (etaint < mineta) {
            etaint = mineta;
            *ico = (*ico == 1) ? 2 : 1;
        }
        eta[istep + 1] = etaint;
        return;
    }
    const double etamaxstep = amp * etamax;
    const double etaextrap = (prod[istep] * eta[istep - 1] - prod[istep - 1] * eta[istep]) / (prod[istep] - prod[istep - 1]);
    eta[istep + 1] = etaextrap;
    if (etaextrap <= 0.0 || etaextrap > etamaxstep) eta[istep + 1] = etamaxstep;
    if (eta[istep + 1] > maxeta && *ico == 1) {
        *ico = 2;
        return;
    }
    if (eta[istep + 1] > maxeta) {
        *ico = 1;
        eta[istep + 1] = maxeta;
    }
}

/* diagonal entry of row q of a CSR matrix (sorted columns) */
static double orc_csr_diag(const long long *rowptr, const int *colind, const double *val, int q)
{
    for (long long k = rowptr[q]; k < rowptr[q + 1]; k++)
        if (colind[k] == q) return val[k];
    return 0.0;
}

/* NRSolver::applyConstraintsToStiffness nrsolver.C:458-500: K_ii *= 1e6 at
 * the directly controlled equations */
static void orc_ddm_stiffness(const long long *rowptr, const int *colind, double *val, int nddm, const int *ddm_eq)
{
    for (int k = 0; k < nddm; k++) {
        const int q = ddm_eq[k] - 1;
        for (long long j = rowptr[q]; j < rowptr[q + 1]; j++)
            if (colind[j] == q) val[j] *= 1.e6;
    }
}

int orc_nonlinear_step(const orc_mesh *m, double *state, double *X, const double *u_presc,
                       const double *u_presc_prev, int initial_guess, const double *RT,
                       double rtolf, int maxiter, int manrmsteps, double lstol, int *out, double *err)
{
    int o4[4];
    const int rc = orc_nonlinear_step_ex(m, state, X, u_presc, u_presc_prev, initial_guess, RT, rtolf, maxiter,
                                         manrmsteps, lstol, 0, 0.8, 2.5, 4.0, 1.0e20, 0, 0, NULL, NULL, 0.0, 0,
                                         o4, err);
    out[0] = o4[0];
    out[1] = o4[1];
    out[2] = o4[2];
    return rc;
}

/* out-of-balance product -(RT - F(X + eta dr)) . dr with the state updated
 * (LineSearchNM::solve linesearch.C:94-118; the eqnmask equations, the
 * ddm ones, are left out of the product) */
static int orc_ls_product(const orc_mesh *m, double *state, const double *rb, const double *dr, double eta,
                          const double *u_presc, const double *RT, const unsigned char *mask, double *X, double *F,
                          double *prod)
{
    for (int i = 0; i < m->neq; i++) X[i] = rb[i] + eta * dr[i];
    if (orc_internal_forces(m, X, u_presc, state, 1, F, NULL, NULL)) return -11;
    double s = 0.0;
    for (int i = 0; i < m->neq; i++)
        if (!mask[i]) s -= ((RT ? RT[i] : 0.0) - F[i]) * dr[i];
    *prod = s;
    return 0;
}

int orc_nonlinear_step_ex(const orc_mesh *m, double *state, double *X, const double *u_presc,
                          const double *u_presc_prev, int initial_guess, const double *RT, double rtolf, int maxiter,
                          int manrmsteps, double lstol, int linesearch, double ls_tol, double ls_amp, double ls_maxeta,
                          double maxinc, int stiffmode, int nddm, const int *ddm_eq, const double *ddm_val,
                          double ddm_factor, int cnr_miniter, int *out, double *err)
{
    orc_nr_opts o;
    orc_nr_opts_default(&o);
    o.initial_guess = initial_guess;
    o.rtolf = rtolf;
    o.maxiter = maxiter;
    o.manrmsteps = manrmsteps;
    o.lstol = lstol;
    o.linesearch = linesearch;
    o.ls_tol = ls_tol;
    o.ls_amp = ls_amp;
    o.ls_maxeta = ls_maxeta;
    o.maxinc = maxinc;
    o.stiffmode = stiffmode;
    o.nddm = nddm;
    o.ddm_eq = ddm_eq;
    o.ddm_val = ddm_val;
    o.ddm_factor = ddm_factor;
    o.cnr_miniter = cnr_miniter;
    return orc_nonlinear_step_opt(m, state, X, u_presc, u_presc_prev, RT, &o, out, err, NULL, 0);
}

void orc_nr_opts_default(orc_nr_opts *o)
{
    memset(o, 0, sizeof *o);
    o->initial_guess = 1;
    o->rtolf = 1.e-3;            /* nrsolver.C:137 */
    o->rtold = 0.0;              /* off, nrsolver.C:138-141 */
    o->maxiter = 100000000;      /* nsmax, nrsolver.C:106 */
    o->lstol = 1.e-5;
    o->ls_tol = 0.8;             /* linesearch.C:46-50 */
    o->ls_amp = 2.5;
    o->ls_maxeta = 4.0;
    o->maxinc = 1.0e20;
}

/* NRSolver::solve nrsolver.C:205-404 and checkConvergence :594-885 (dof
 * groups u, v, w), LineSearchNM::solve linesearch.C:54-160 */
int orc_nonlinear_step_opt(const orc_mesh *m, double *state, double *X, const double *u_presc,
                           const double *u_presc_prev, const double *RT, const orc_nr_opts *o, int *out, double *err,
                           double *errhist, int errcap)
{
    const int tangent = (o->stiffmode == 2 || o->stiffmode == 1 || o->stiffmode == 3) ? 0 : 1;
    const int nddm = o->nddm;
    const int *ddm_eq = o->ddm_eq;
    /* NRSolver's constrained Newton (nrsolver.C:183-185, 320-326, 612-623):
     * from iteration cnr_miniter + 1 on, a correction is halved when the
     * squared norm of the per-dof-group force errors grew */
    double fe_new[3] = {0, 0, 0}, fe_old[3] = {0, 0, 0};
    const long long nnz = orc_csr_count(m);
    long long *rowptr = (long long *)malloc(sizeof(long long) * ((size_t)m->neq + 1));
    int *colind = (int *)malloc(sizeof(int) * (size_t)nnz);
    double *val = (double *)malloc(sizeof(double) * (size_t)nnz);
    double *F = (double *)malloc(sizeof(double) * ((size_t)m->neq + 1));
    double *rhs = (double *)malloc(sizeof(double) * ((size_t)m->neq + 1));
    double *ddX = (double *)calloc((size_t)m->neq + 1, sizeof(double));
    int *dofid = (int *)malloc(sizeof(int) * ((size_t)m->neq + 1));
    unsigned char *mask = (unsigned char *)calloc((size_t)m->neq + 1, 1);
    int inuse[3] = {0, 0, 0}, rc = 0;
    for (long long i = 0; i < 3LL * m->nnode; i++)
        if (m->eqn[i] > 0) { dofid[m->eqn[i] - 1] = (int)(i % 3); inuse[i % 3]++; }
    for (int k = 0; k < nddm; k++) mask[ddm_eq[k] - 1] = 1;
    double *rb = (double *)malloc(sizeof(double) * ((size_t)m->neq + 1));
    orc_csr_fill(m, rowptr, colind);
    out[0] = out[1] = out[2] = out[3] = 0;
    /* LineSearchNM::initializeFrom clamps (linesearch.C:247-283) */
    double ls_tol = o->ls_tol, ls_amp = o->ls_amp, ls_maxeta = o->ls_maxeta;
    if (ls_tol < 0.6) ls_tol = 0.6;
    if (ls_tol > 0.95) ls_tol = 0.95;
    if (ls_amp < 1.0) ls_amp = 1.0;
    if (ls_amp > 10.0) ls_amp = 10.0;
    if (ls_maxeta < 1.5) ls_maxeta = 1.5;
    if (ls_maxeta > 15.0) ls_maxeta = 15.0;
    if (o->initial_guess && m->npeq > 0) {
        double *du = (double *)malloc(sizeof(double) * (size_t)m->npeq);
        for (int i = 0; i < m->npeq; i++) du[i] = u_presc[i] - (u_presc_prev ? u_presc_prev[i] : 0.0);
        rc = orc_extrapolated_forces(m, state, du, rhs);
        free(du);
        if (rc) { rc = -12; goto out; }
        double nrm = 0.0;
        for (int i = 0; i < m->neq; i++) { rhs[i] = -rhs[i]; nrm += rhs[i] * rhs[i]; }
        if (nrm > 0.0) {
            if (orc_assemble_stiffness(m, rowptr, colind, val, state, tangent)) { rc = -10; goto out; }
            int it = 0;
            double res = 0.0;
            orc_pcg(m->neq, rowptr, colind, val, rhs, ddX, o->lstol, 1000000, 1, &it, &res);
            out[1] += it;
            for (int i = 0; i < m->neq; i++) { X[i] += ddX[i]; ddX[i] = 0.0; }
        }
    }
    if (orc_assemble_stiffness(m, rowptr, colind, val, state, tangent)) { rc = -10; goto out; }
    orc_ddm_stiffness(rowptr, colind, val, nddm, ddm_eq);
    for (int nite = 0;; nite++) {
        double ebe[3];
        if (orc_internal_forces(m, X, u_presc, state, 1, F, NULL, ebe)) { rc = -11; goto out; }
        double s_r[3] = {0, 0, 0}, s_t[3] = {0, 0, 0}, s_dd[3] = {0, 0, 0}, s_x[3] = {0, 0, 0};
        for (int i = 0; i < m->neq; i++) rhs[i] = (RT ? RT[i] : 0.0) - F[i];
        /* NRSolver::applyConstraintsToLoadIncrement nrsolver.C:502-568 */
        for (int k = 0; k < nddm; k++) {
            const int q = ddm_eq[k] - 1;
            rhs[q] = nite == 0 ? orc_csr_diag(rowptr, colind, val, q) * o->ddm_val[k] * o->ddm_factor : 0.0;
        }
        for (int i = 0; i < m->neq; i++) {
            const double rt = RT ? RT[i] : 0.0;
            s_r[dofid[i]] += rhs[i] * rhs[i];
            s_t[dofid[i]] += rt * rt;
            s_dd[dofid[i]] += ddX[i] * ddX[i];
            s_x[dofid[i]] += X[i] * X[i];
        }
        int converged = 1, outofrange = 0;
        double mx = 0.0;
        if (o->cnr_miniter) {
            for (int d = 0; d < 3; d++) { fe_old[d] = fe_new[d]; fe_new[d] = 0.0; }
        }
        double *eh = (errhist && nite < errcap) ? errhist + 6 * (size_t)nite : NULL;
        for (int d = 0; d < 3; d++) {
            if (eh) { eh[d] = -1.0; eh[3 + d] = -1.0; }
            if (!inuse[d]) continue;
            if (o->rtolf > 0.0) {
                /* force error of the group (nrsolver.C:754-782) */
                const double den = s_t[d] + ebe[d];
                const double fe = den >= 1.e-6 ? sqrt(s_r[d] / den) : sqrt(s_r[d]);
                fe_new[d] = fe;
                if (isnan(fe) || fe > o->rtolf * 1.e20) outofrange = 1;
                if (fe > o->rtolf) converged = 0;
                if (fe > 0.0 && nite == 0) converged = 0;
                if (fe > mx) mx = fe;
                if (eh) eh[d] = fe;
            }
            if (o->rtold > 0.0) {
                /* displacement error of the group (nrsolver.C:786-806) */
                const double de = s_x[d] > 1.e-6 ? sqrt(s_dd[d] / s_x[d]) : sqrt(s_dd[d]);
                if (isnan(de) || de > o->rtold * 1.e20) outofrange = 1;
                if (de > o->rtold) converged = 0;
                if (eh) eh[3 + d] = de;
            }
        }
        *err = mx;
        out[0] = nite;
        if (outofrange) { out[2] = -2; break; }
        if (converged && nite >= o->miniter) { out[2] = 1; break; }
        if (nite >= o->maxiter) { out[2] = -1; break; }
        if (nite > 0 && o->manrmsteps > 0 && nite % o->manrmsteps == 0) {
            if (orc_assemble_stiffness(m, rowptr, colind, val, state, tangent)) { rc = -10; goto out; }
            orc_ddm_stiffness(rowptr, colind, val, nddm, ddm_eq);
        }
        int it = 0;
        double res = 0.0;
        orc_pcg(m->neq, rowptr, colind, val, rhs, ddX, o->lstol, 1000000, 1, &it, &res);
        out[1] += it;
        const int do_ls = o->linesearch && nite > 0;
        if (!do_ls && o->cnr_miniter && nite > o->cnr_miniter) {
            const double a = fe_new[0] * fe_new[0] + fe_new[1] * fe_new[1] + fe_new[2] * fe_new[2];
            const double b = fe_old[0] * fe_old[0] + fe_old[1] * fe_old[1] + fe_old[2] * fe_old[2];
            if (a > b) for (int i = 0; i < m->neq; i++) ddX[i] *= 0.5;
        }
        double mxi = 0.0;
        if (o->ls_keep_point || !do_ls) {
            /* maxIncAllowed (nrsolver.C:332-345).  The reference caps after
             * the line search; ls_keep_point caps before it so that the
             * search runs along the capped direction */
            for (int i = 0; i < m->neq; i++) if (fabs(ddX[i]) > mxi) mxi = fabs(ddX[i]);
            if (mxi > o->maxinc) for (int i = 0; i < m->neq; i++) ddX[i] *= o->maxinc / mxi;
        }
        if (do_ls) {
            /* LineSearchNM::solve linesearch.C:54-160 with eqnmask = the ddm
             * equations.  It moves X itself: to X_start + eta dr when the
             * search meets its tolerance (dr scaled by eta), else to
             * X_start + dr, with the state of that point. */
            double prod[12], eta[12];
            for (int i = 0; i < m->neq; i++) rb[i] = X[i];
            double s0 = 0.0;
            for (int i = 0; i < m->neq; i++)
                if (!mask[i]) s0 -= ((RT ? RT[i] : 0.0) - F[i]) * ddX[i];
            double etav = 1.0;
            int ok = 0;
            if (s0 < 0.0) {
                const int max_iter = 10;
                int ico = 0, ils;
                prod[1] = 1.0;
                eta[1] = 0.0;
                eta[2] = 1.0;
                for (ils = 2; ils <= max_iter; ils++) {
                    double si = 0.0;
                    if (orc_ls_product(m, state, rb, ddX, eta[ils], u_presc, RT, mask, X, F, &si)) {
                        rc = -11;
                        goto out;
                    }
                    si /= s0;
                    prod[ils] = si;
                    if (fabs(si) < ls_tol) { etav = eta[ils]; ok = 1; break; }
                    orc_ls_search(ils, prod, eta, ls_amp, ls_maxeta, 0.2, &ico);
                    if (ico == 2) break;
                }
            }
            if (!ok) etav = 1.0;
            if (etav != 1.0) out[3]++;
            for (int i = 0; i < m->neq; i++) X[i] = rb[i] + etav * ddX[i];
            if (ok) for (int i = 0; i < m->neq; i++) ddX[i] *= etav;
            if (!o->ls_keep_point) {
                /* reference order: the cap of the (scaled) correction after
                 * the search (nrsolver.C:332-345), then X.add(ddX) (:349),
                 * which passes the line search's step a second time */
                for (int i = 0; i < m->neq; i++) if (fabs(ddX[i]) > mxi) mxi = fabs(ddX[i]);
                if (mxi > o->maxinc) for (int i = 0; i < m->neq; i++) ddX[i] *= o->maxinc / mxi;
                for (int i = 0; i < m->neq; i++) X[i] += ddX[i];
            }
        } else {
            for (int i = 0; i < m->neq; i++) X[i] += ddX[i];
        }
    }
out:
    free(rowptr); free(colind); free(val); free(F); free(rhs); free(ddX); free(dofid); free(rb); free(mask);
    return rc;
}

/* ------------------------------------------------------------------ */
/* LinearStatic::solveYourselfAt linearstatic.C:177-252: build K, assemble,
 * rhs = f_ext - f_int(prescribed values), CG from x=0; then reactions
 * StructuralEngngModel::computeReaction structengngmodel.C:183-201:
 * R = f_int(prescribed numbering) - f_ext(prescribed numbering).       */
/* ------------------------------------------------------------------ */
int orc_linear_static(const orc_mesh *m, const double *f_ext, const double *f_ext_presc,
                      const double *u_presc, double tol, int maxit,
                      double *u, double *reactions, int *iters, double *resid)
{
    const long long nnz = orc_csr_count(m);
    long long *rowptr = (long long *)malloc(sizeof(long long) * ((size_t)m->neq + 1));
    int *colind = (int *)malloc(sizeof(int) * (size_t)nnz);
    double *val = (double *)malloc(sizeof(double) * (size_t)nnz);
    double *rhs = (double *)malloc(sizeof(double) * ((size_t)m->neq + 1));
    double *fp = (double *)malloc(sizeof(double) * ((size_t)m->npeq + 1));
    int rc = 0;
    orc_csr_fill(m, rowptr, colind);
    if (orc_assemble_stiffness(m, rowptr, colind, val, NULL, 0)) { rc = -10; goto out; }
    for (int i = 0; i < m->neq; i++) u[i] = 0.0;
    orc_internal_forces(m, u, u_presc, NULL, 0, rhs, NULL, NULL);
    for (int i = 0; i < m->neq; i++) rhs[i] = f_ext[i] - rhs[i];
    rc = orc_pcg(m->neq, rowptr, colind, val, rhs, u, tol, maxit, 1, iters, resid);
    if (reactions) {
        orc_internal_forces(m, u, u_presc, NULL, 0, NULL, fp, NULL);
        for (int i = 0; i < m->npeq; i++) reactions[i] = fp[i] - (f_ext_presc ? f_ext_presc[i] : 0.0);
    }
out:
    free(rowptr); free(colind); free(val); free(rhs); free(fp);
    return rc;
}

/* ------------------------------------------------------------------ */
/* Profile (skyline) U^T D U factorization and solve, the reference's
 * direct solver: Skyline::factorized skyline.C:429-497 (Crout, column by
 * column) and Skyline::backSubstitutionWith skyline.C:248-290, restated on
 * 0-based column profiles: column j holds rows first[j]..j, stored from the
 * first row down to the diagonal.  Input is the full symmetric CSR.  The
 * inner sums run over ascending rows as in the reference.  Returns 0, or
 * -1 on a zero pivot, -2 on allocation failure.                        */
/* ------------------------------------------------------------------ */
int orc_skyline_solve(int n, const long long *rowptr, const int *colind, const double *val, const double *b,
                      double *x, long long *nwk_out)
{
    int *first = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    long long *cs = (long long *)malloc(sizeof(long long) * ((size_t)n + 1));
    if (!first || !cs) { free(first); free(cs); return -2; }
    for (int j = 0; j < n; j++) {
        int f = j;
        for (long long k = rowptr[j]; k < rowptr[j + 1]; k++)
            if (colind[k] < f) f = colind[k];
        first[j] = f;
    }
    cs[0] = 0;
    for (int j = 0; j < n; j++) cs[j + 1] = cs[j] + (j - first[j] + 1);
    if (nwk_out) *nwk_out = cs[n];
    double *a = (double *)calloc((size_t)cs[n] + 1, sizeof(double));
    double *y = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    if (!a || !y) { free(first); free(cs); free(a); free(y); return -2; }
#define SKY(i, j) a[cs[(j)] + ((i) - first[(j)])]   /* entry (i, j), first[j] <= i <= j */
    /* upper triangle of row i = column entries (i, j), j >= i */
    for (int i = 0; i < n; i++)
        for (long long k = rowptr[i]; k < rowptr[i + 1]; k++) {
            const int j = colind[k];
            if (j >= i) SKY(i, j) = val[k];
        }
    int rc = 0;
    for (int k = 1; k < n; k++) {
        const int fk = first[k];
        for (int i = fk + 1; i < k; i++) {
            const int fi = first[i];
            const int lo = fi > fk ? fi : fk;
            double s = 0.0;
            for (int m = lo; m < i; m++) s += SKY(m, k) * SKY(m, i);
            SKY(i, k) -= s;
        }
        double s = 0.0;
        for (int i = fk; i < k; i++) {
            const double g = SKY(i, k);
            SKY(i, k) /= SKY(i, i);
            s += SKY(i, k) * g;
        }
        SKY(k, k) -= s;
        if (SKY(k, k) == 0.0) { rc = -1; break; }
    }
    if (!rc && n > 0 && SKY(0, 0) == 0.0) rc = -1;
    if (!rc) {
        for (int i = 0; i < n; i++) y[i] = b[i];
        for (int k = 1; k < n; k++) {
            double s = 0.0;
            for (int m = first[k]; m < k; m++) s += SKY(m, k) * y[m];
            y[k] -= s;
        }
        for (int k = 0; k < n; k++) y[k] /= SKY(k, k);
        for (int k = n - 1; k >= 0; k--) {
            x[k] = y[k];
            for (int m = first[k]; m < k; m++) y[m] -= SKY(m, k) * x[k];
        }
    }
#undef SKY
    free(first); free(cs); free(a); free(y);
    return rc;
}

// Device form of MisesMat (isotropic J2 plasticity with linear isotropic
// hardening, no damage), shared by the element kernels and the surrogate's
// training-data generator.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace hipfem {

// ----------------------------------------------------------------- MisesMat
// performPlasticityReturn misesmat.C:181-263 (3D branch :221-256), helpers
// structuralmaterial.C:1489-1600; no damage (omega_crit = 0).
struct MisesOut {
    double sig[6], epsp[6], sdev[6], kappa;
};

__device__ __forceinline__ void mises_return(double E, double nu, double sig0, double H, const double eps[6],
                                             const double epsp_n[6], double kappa_n, MisesOut &o)
{
    const double G = E / (2.0 * (1. + nu));
    const double K = E / (3.0 * (1. - 2. * nu));
    double el[6];
#pragma unroll
    for (int i = 0; i < 6; i++) {
        o.epsp[i] = epsp_n[i];
        el[i] = eps[i] - epsp_n[i];
    }
    const double mean = (el[0] + el[1] + el[2]) / 3.0;
    el[0] -= mean;
    el[1] -= mean;
    el[2] -= mean;
    double s[6] = {2. * G * el[0], 2. * G * el[1], 2. * G * el[2], G * el[3], G * el[4], G * el[5]};
    const double pvol = 3 * K * mean;
#pragma unroll
    for (int i = 0; i < 6; i++) o.sdev[i] = s[i];
    const double trialS = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2] + 2. * s[3] * s[3] + 2. * s[4] * s[4] +
                               2. * s[5] * s[5]);
    double kappa = kappa_n;
    const double yieldValue = sqrt(3. / 2.) * trialS - (sig0 + H * kappa);
    if (yieldValue > 0.) {
        const double dKappa = yieldValue / (H + 3. * G);
        kappa += dKappa;
        const double f = sqrt(3. / 2.) * dKappa / trialS;
        o.epsp[0] += f * s[0];
        o.epsp[1] += f * s[1];
        o.epsp[2] += f * s[2];
        o.epsp[3] += f * (2. * s[3]);
        o.epsp[4] += f * (2. * s[4]);
        o.epsp[5] += f * (2. * s[5]);
        const double sc = 1. - sqrt(6.) * G * dKappa / trialS;
#pragma unroll
        for (int i = 0; i < 6; i++) s[i] *= sc;
    }
    s[0] += pvol;
    s[1] += pvol;
    s[2] += pvol;
#pragma unroll
    for (int i = 0; i < 6; i++) o.sig[i] = s[i];
    o.kappa = kappa;
}

// give3dMaterialStiffnessMatrix misesmat.C:496-544 (TangentStiffness, no damage)
__device__ __forceinline__ void mises_tangent(double E, double nu, double sig0, double H, double kappa_n,
                                              double kappa_t, const double s[6], double D[36])
{
    const double G = E / (2.0 * (1. + nu));
    const double K = E / (3.0 * (1. - 2. * nu));
    const double idev[36] = {2. / 3.,  -1. / 3., -1. / 3., 0., 0.,  0.,  -1. / 3., 2. / 3., -1. / 3.,
                             0.,       0.,       0.,       -1. / 3., -1. / 3., 2. / 3., 0., 0.,  0.,
                             0.,       0.,       0.,       0.5,      0.,       0.,       0., 0.,  0.,
                             0.,       0.5,      0.,       0.,       0.,       0.,       0., 0.,  0.5};
#pragma unroll
    for (int i = 0; i < 36; i++) {
        const int r = i / 6, c = i % 6;
        D[i] = 2 * G * idev[i] + ((r < 3 && c < 3) ? K : 0.0);
    }
    const double dKappa = kappa_t - kappa_n;
    if (dKappa <= 0.0) return;
    const double sigmaY = sig0 + H * kappa_n;
    const double trialS = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2] + 2. * s[3] * s[3] + 2. * s[4] * s[4] +
                               2. * s[5] * s[5]);
    const double factor = -2. * sqrt(6.) * G * G / trialS;
    const double factor1 = factor * sigmaY / ((H + 3. * G) * trialS * trialS);
    const double factor2 = factor * dKappa;
#pragma unroll
    for (int i = 0; i < 36; i++) D[i] += factor1 * (s[i / 6] * s[i % 6]);
#pragma unroll
    for (int i = 0; i < 36; i++) D[i] += factor2 * idev[i];
}

// The same tangent in the form D = c I_dev + K 1x1 - m m^T (I_dev as in
// mises_tangent; 1 the Voigt unit in the three normal rows): c = 2G +
// factor2 and m = sqrt(-factor1) s (factor1 <= 0), and c = 2G, m = 0 on the
// elastic branch.  Returns K.
__device__ __forceinline__ double mises_tangent_parts(double E, double nu, double sig0, double H, double kappa_n,
                                                      double kappa_t, const double s[6], double &c, double m[6])
{
    const double G = E / (2.0 * (1. + nu));
    const double K = E / (3.0 * (1. - 2. * nu));
    c = 2. * G;
#pragma unroll
    for (int i = 0; i < 6; i++) m[i] = 0.0;
    const double dKappa = kappa_t - kappa_n;
    if (dKappa <= 0.0) return K;
    const double sigmaY = sig0 + H * kappa_n;
    const double trialS = sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2] + 2. * s[3] * s[3] + 2. * s[4] * s[4] +
                               2. * s[5] * s[5]);
    const double factor = -2. * sqrt(6.) * G * G / trialS;
    const double factor1 = factor * sigmaY / ((H + 3. * G) * trialS * trialS);
    c += factor * dKappa;
    const double sc = sqrt(-factor1);
#pragma unroll
    for (int i = 0; i < 6; i++) m[i] = sc * s[i];
    return K;
}

// Elastic trial stress D_e (eps - epsp_n) (Voigt, engineering shear) and
// the entries of the upper triangle of D_e (row-major, k = 0..20).
__device__ __forceinline__ void elastic_trial(double E, double nu, const double el[6], double sig[6])
{
    const double G = E / (2.0 * (1. + nu));
    const double K = E / (3.0 * (1. - 2. * nu));
    const double mean = (el[0] + el[1] + el[2]) / 3.0;
    for (int i = 0; i < 3; i++) sig[i] = 2. * G * (el[i] - mean) + 3. * K * mean;
    for (int i = 3; i < 6; i++) sig[i] = G * el[i];
}

__device__ __forceinline__ double elastic_upper(double E, double nu, int k)
{
    const double G = E / (2.0 * (1. + nu));
    const double K = E / (3.0 * (1. - 2. * nu));
    int r = 0, c = 0, t = k;
    for (r = 0; r < 6; r++) {
        if (t < 6 - r) {
            c = r + t;
            break;
        }
        t -= 6 - r;
    }
    if (r < 3 && c < 3) return (r == c ? 2. * G * 2. / 3. : -2. * G / 3.) + K;
    return r == c ? G : 0.0;
}

}  // namespace hipfem

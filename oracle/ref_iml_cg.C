/*
 * ref_iml_cg.C -- TEST INFRASTRUCTURE ONLY (part of the oracle).
 *
 * Compiles the reference's own conjugate-gradient template, iml/cg.h, as it
 * stands in the reference tree (the Makefile adds the reference's iml/
 * directory to the include path; nothing of it is copied here).  cg.h has no
 * includes of its own: it is a template over a Matrix, a Vector, a
 * Preconditioner and a Real.  IMLSolver::solve (src/core/iml/imlsolver.C:102-146)
 * instantiates it with SparseMtrx, FloatArray, DiagPreconditioner and double.
 * Those three classes need the reference's CMake-generated headers, so this
 * file instantiates cg.h with small classes of its own that perform the same
 * arithmetic in the same order:
 *   - Vec: the IML operators of FloatArray (floatarray.C:958-1030, element
 *     loops of beScaled :208, add :218, subtract :320, beDifferenceOf :403;
 *     dot = std::inner_product in index order, dotProduct :524; norm =
 *     sqrt of that, computeNorm :861-869);
 *   - SymCompColMat / CompColMat: SparseMtrx::operator* (sparsemtrx.h:289)
 *     through SymCompCol::times (symcompcol.C:201-218, lower triangle per
 *     column, diagonal first) or CompCol::times (compcol.C:128-142);
 *   - DiagPrec: DiagPreconditioner::init / solve (diagpre.C:41-68).
 *
 * The C entry point records every norm() that CG takes: the first is ||b||,
 * the rest are ||r|| of the initial residual and of every iteration, which
 * gives the reference's residual history without touching cg.h.
 */
#include <cmath>
#include <cstdint>
#include <numeric>
#include <vector>

namespace refcg {

struct Vec {
    std::vector<double> v;
    Vec() = default;
    explicit Vec(int n) : v((size_t)n, 0.0) {}
    double &operator()(int i) { return v[(size_t)i]; }
    double operator()(int i) const { return v[(size_t)i]; }
    int size() const { return (int)v.size(); }
};

static std::vector<double> *g_norms = nullptr;

inline double dot(const Vec &a, const Vec &b) { return std::inner_product(a.v.begin(), a.v.end(), b.v.begin(), 0.); }

inline double norm(const Vec &a)
{
    const double s = std::sqrt(dot(a, a));
    if (g_norms) g_norms->push_back(s);
    return s;
}

inline Vec operator*(const double &a, const Vec &x)
{
    Vec r(x.size());
    for (size_t i = 0; i < x.v.size(); i++) r.v[i] = a * x.v[i];
    return r;
}

inline Vec operator+(const Vec &x, const Vec &y)
{
    Vec r(x);
    for (size_t i = 0; i < r.v.size(); i++) r.v[i] = r.v[i] + y.v[i];
    return r;
}

inline Vec operator-(const Vec &x, const Vec &y)
{
    Vec r(x.size());
    for (size_t i = 0; i < x.v.size(); i++) r.v[i] = x.v[i] - y.v[i];
    return r;
}

inline Vec &operator+=(Vec &x, const Vec &y)
{
    for (size_t i = 0; i < x.v.size(); i++) x.v[i] = x.v[i] + y.v[i];
    return x;
}

inline Vec &operator-=(Vec &x, const Vec &y)
{
    for (size_t i = 0; i < x.v.size(); i++) x.v[i] = x.v[i] - y.v[i];
    return x;
}

// lower triangle by columns, rows ascending, diagonal first in every column
struct SymCompColMat {
    int n = 0;
    std::vector<int64_t> colptr;
    std::vector<int> rowind;
    std::vector<double> val;
    Vec operator*(const Vec &x) const
    {
        Vec y(n);
        for (int j = 0; j < n; j++) {
            const double rhs = x.v[(size_t)j];
            double sum = 0.0;
            for (int64_t t = colptr[(size_t)j] + 1; t < colptr[(size_t)j + 1]; t++) {
                y.v[(size_t)rowind[(size_t)t]] += val[(size_t)t] * rhs;
                sum += val[(size_t)t] * x.v[(size_t)rowind[(size_t)t]];
            }
            y.v[(size_t)j] += sum + val[(size_t)colptr[(size_t)j]] * rhs;
        }
        return y;
    }
    double diag(int i) const { return val[(size_t)colptr[(size_t)i]]; }
};

struct CompColMat {
    int n = 0;
    std::vector<int64_t> colptr;
    std::vector<int> rowind;
    std::vector<double> val;
    Vec operator*(const Vec &x) const
    {
        Vec y(n);
        for (int j = 0; j < n; j++) {
            const double rhs = x.v[(size_t)j];
            for (int64_t t = colptr[(size_t)j]; t < colptr[(size_t)j + 1]; t++)
                y.v[(size_t)rowind[(size_t)t]] += val[(size_t)t] * rhs;
        }
        return y;
    }
};

struct DiagPrec {
    std::vector<double> d;
    Vec solve(const Vec &x) const
    {
        Vec y(x.size());
        for (size_t i = 0; i < x.v.size(); i++) y.v[i] = x.v[i] * d[i];
        return y;
    }
};

#include "cg.h"

}  // namespace refcg

extern "C" {

/*
 * CG of iml/cg.h on the symmetric matrix given as a full CSR (0-based,
 * columns sorted per row).  storage 0: SymCompCol order, 1: CompCol order.
 * precond 1: diagonal (zero diagonal -> return -1), 0: identity.  x holds
 * the initial guess on entry and the solution on exit.  Returns cg.h's flag
 * (0 converged, 1 not within maxit); *iters and *tol receive its max_iter
 * and tol outputs.  norms (capacity cap, may be NULL) receives every norm()
 * CG takes: ||b||, ||r_0||, ||r_1|| ...; *nnorms their count.
 */
int ref_iml_cg(int n, const int64_t *rowptr, const int *colind, const double *val, const double *b, double *x,
               int maxit, double tol_in, int storage, int precond, int *iters, double *tol_out, double *norms, int cap,
               int *nnorms)
{
    using namespace refcg;
    SymCompColMat S;
    CompColMat C;
    DiagPrec M;
    M.d.assign((size_t)n, 1.0);
    if (storage == 0) {
        S.n = n;
        S.colptr.assign((size_t)n + 1, 0);
        for (int j = 0; j < n; j++) {
            // column j of the lower triangle = row j's entries with col >= j
            for (int64_t k = rowptr[j]; k < rowptr[j + 1]; k++)
                if (colind[k] >= j) {
                    S.rowind.push_back(colind[k]);
                    S.val.push_back(val[k]);
                }
            S.colptr[(size_t)j + 1] = (int64_t)S.rowind.size();
            if (S.colptr[(size_t)j + 1] == S.colptr[(size_t)j] || S.rowind[(size_t)S.colptr[(size_t)j]] != j)
                return -2;  // the diagonal must be stored
        }
    } else {
        // column j = row j for a symmetric pattern and values
        C.n = n;
        C.colptr.assign(rowptr, rowptr + n + 1);
        C.rowind.assign(colind, colind + rowptr[n]);
        C.val.assign(val, val + rowptr[n]);
    }
    if (precond == 1)
        for (int i = 0; i < n; i++) {
            double d = 0.0;
            for (int64_t k = rowptr[i]; k < rowptr[i + 1]; k++)
                if (colind[k] == i) d = val[k];
            if (d == 0) return -1;
            M.d[(size_t)i] = 1. / d;
        }
    Vec xv(n), bv(n);
    for (int i = 0; i < n; i++) {
        xv.v[(size_t)i] = x[i];
        bv.v[(size_t)i] = b[i];
    }
    std::vector<double> rec;
    g_norms = &rec;
    int mi = maxit;
    double t = tol_in;
    const int flag = storage == 0 ? CG(S, xv, bv, M, mi, t) : CG(C, xv, bv, M, mi, t);
    g_norms = nullptr;
    for (int i = 0; i < n; i++) x[i] = xv.v[(size_t)i];
    *iters = mi;
    *tol_out = t;
    if (nnorms) *nnorms = (int)rec.size();
    if (norms)
        for (int i = 0; i < (int)rec.size() && i < cap; i++) norms[i] = rec[(size_t)i];
    return flag;
}

}  // extern "C"

// Minimal host-side types with the reference's names and conventions, so
// the plug-in classes below compile against the same signatures they would
// have inside the reference (FloatArray / IntArray / FloatMatrix with 1-based
// at(), the SparseMtrxType / LinSystSolverType / ConvergedReason enums,
// OOFEM_ERROR raising RuntimeException).
//
// Reference declarations mirrored: src/core/floatarray.h, intarray.h,
// floatmatrix.h (column-major, 1-based at()), sparsemtrxtype.h:42-54,
// linsystsolvertype.h:47-57, convergedreason.h:41-50, error.h:78-80.
#pragma once

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace oofem {

class RuntimeException : public std::runtime_error {
public:
    explicit RuntimeException(const std::string &m) : std::runtime_error(m) {}
};

[[noreturn]] inline void oofem_error(const char *func, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    throw RuntimeException(std::string(func) + ": " + buf);
}
#define OOFEM_ERROR(...) ::oofem::oofem_error(__func__, __VA_ARGS__)

// sparsemtrxtype.h:42-54; the existing values are deck integers and stay put,
// SMT_HipCSR is appended.
enum SparseMtrxType {
    SMT_Skyline, SMT_SkylineU, SMT_CompCol, SMT_DynCompCol, SMT_SymCompCol, SMT_DynCompRow,
    SMT_SpoolesMtrx, SMT_PetscMtrx, SMT_DSS_sym_LDL, SMT_DSS_sym_LL, SMT_DSS_unsym_LU,
    SMT_HipCSR  // = 11
};

// linsystsolvertype.h:47-57; ST_HipPCG appended.
enum LinSystSolverType {
    ST_Direct = 0, ST_IML = 1, ST_Spooles = 2, ST_Petsc = 3, ST_DSS = 4, ST_Feti = 5, ST_MKLPardiso = 6,
    ST_SuperLU_MT = 7, ST_PardisoProjectOrg = 8,
    ST_HipPCG = 9
};

// convergedreason.h:41-50
enum ConvergedReason { CR_UNKNOWN = 0, CR_CONVERGED = 1, CR_DIVERGED_ITS = -1, CR_DIVERGED_TOL = -2, CR_FAILED = -99 };

class FloatArray {
    std::vector<double> v;

public:
    FloatArray(int n = 0) : v((size_t)n, 0.0) {}
    FloatArray(std::initializer_list<double> l) : v(l) {}
    int giveSize() const { return (int)v.size(); }
    bool isNotEmpty() const { return !v.empty(); }
    void resize(int n) { v.assign((size_t)n, 0.0); }
    void resizeWithValues(int n) { v.resize((size_t)n, 0.0); }
    void zero() { std::fill(v.begin(), v.end(), 0.0); }
    void clear() { v.clear(); }
    double &at(int i) { return v[(size_t)i - 1]; }
    double at(int i) const { return v[(size_t)i - 1]; }
    double &operator[](int i) { return v[(size_t)i]; }
    double operator[](int i) const { return v[(size_t)i]; }
    double &operator()(int i) { return v[(size_t)i]; }
    double *givePointer() { return v.data(); }
    const double *givePointer() const { return v.data(); }
    void add(const FloatArray &b)
    {
        if (v.empty()) { v = b.v; return; }
        for (size_t i = 0; i < v.size(); i++) v[i] += b.v[i];
    }
    void add(double f, const FloatArray &b)
    {
        if (v.empty()) v.assign(b.v.size(), 0.0);
        for (size_t i = 0; i < v.size(); i++) v[i] += f * b.v[i];
    }
    void subtract(const FloatArray &b)
    {
        if (v.empty()) v.assign(b.v.size(), 0.0);
        for (size_t i = 0; i < v.size(); i++) v[i] -= b.v[i];
    }
    void beDifferenceOf(const FloatArray &a, const FloatArray &b)
    {
        v.resize(a.v.size());
        for (size_t i = 0; i < v.size(); i++) v[i] = a.v[i] - b.v[i];
    }
    void times(double f)
    {
        for (double &x : v) x *= f;
    }
    double computeSquaredNorm() const
    {
        double s = 0.0;
        for (double x : v) s += x * x;
        return s;
    }
    double computeNorm() const { return std::sqrt(computeSquaredNorm()); }
    std::vector<double>::iterator begin() { return v.begin(); }
    std::vector<double>::iterator end() { return v.end(); }
    std::vector<double>::const_iterator begin() const { return v.begin(); }
    std::vector<double>::const_iterator end() const { return v.end(); }
};

class IntArray {
    std::vector<int> v;

public:
    IntArray(int n = 0) : v((size_t)n, 0) {}
    IntArray(std::initializer_list<int> l) : v(l) {}
    int giveSize() const { return (int)v.size(); }
    void resize(int n) { v.assign((size_t)n, 0); }
    int &at(int i) { return v[(size_t)i - 1]; }
    int at(int i) const { return v[(size_t)i - 1]; }
    int &operator[](int i) { return v[(size_t)i]; }
    int operator[](int i) const { return v[(size_t)i]; }
    int *givePointer() { return v.data(); }
    const int *givePointer() const { return v.data(); }
    void clear() { v.clear(); }
    void followedBy(int b) { v.push_back(b); }
    std::vector<int>::iterator begin() { return v.begin(); }
    std::vector<int>::iterator end() { return v.end(); }
    std::vector<int>::const_iterator begin() const { return v.begin(); }
    std::vector<int>::const_iterator end() const { return v.end(); }
};

// column-major like the reference's FloatMatrix
class FloatMatrix {
    int nr = 0, nc = 0;
    std::vector<double> v;

public:
    FloatMatrix(int r = 0, int c = 0) : nr(r), nc(c), v((size_t)r * c, 0.0) {}
    int giveNumberOfRows() const { return nr; }
    int giveNumberOfColumns() const { return nc; }
    bool isNotEmpty() const { return nr > 0 && nc > 0; }
    void resize(int r, int c)
    {
        nr = r;
        nc = c;
        v.assign((size_t)r * c, 0.0);
    }
    void zero() { std::fill(v.begin(), v.end(), 0.0); }
    double &at(int i, int j) { return v[(size_t)(j - 1) * nr + (i - 1)]; }
    double at(int i, int j) const { return v[(size_t)(j - 1) * nr + (i - 1)]; }
    const double *givePointer() const { return v.data(); }
};

// src/core/unknownnumberingscheme.h: the default numbering is the only one
// the path uses (free equations, 1-based)
class UnknownNumberingScheme {
public:
    virtual ~UnknownNumberingScheme() {}
    virtual bool isDefault() const { return true; }
};
class EModelDefaultEquationNumbering : public UnknownNumberingScheme {};

}  // namespace oofem

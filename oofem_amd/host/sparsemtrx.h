// SparseMtrx and SparseLinearSystemNM: the two abstract plug-in interfaces of
// the reference (src/core/sparsemtrx.h:60-281, src/core/sparselinsystemnm.h:
// 56-97), restated with the same virtual signatures so that HipCSRSparseMtrx
// and HipPCGSolver here are the classes a maintainer drops into the
// reference tree (see the integration notes for the include swap).
#pragma once

#include <memory>

#include "inputrecord.h"
#include "oofem_types.h"

namespace oofem {

class EngngModel;
class Domain;

class SparseMtrx {
public:
    typedef long SparseMtrxVersionType;

protected:
    int nRows;
    int nColumns;
    SparseMtrxVersionType version;

public:
    SparseMtrx(int n = 0, int m = 0) : nRows(n), nColumns(m), version(0) {}
    virtual ~SparseMtrx() {}

    SparseMtrxVersionType giveVersion() { return this->version; }
    void checkBounds(int i, int j) const
    {
        if (i <= 0) OOFEM_ERROR("matrix error on rows : %d <= 0", i);
        if (j <= 0) OOFEM_ERROR("matrix error on columns : %d <= 0", j);
        if (i > nRows) OOFEM_ERROR("matrix error on rows : %d => %d", i, nRows);
        if (j > nColumns) OOFEM_ERROR("matrix error on columns : %d => %d", j, nColumns);
    }
    int giveNumberOfRows() const { return nRows; }
    int giveNumberOfColumns() const { return nColumns; }
    bool isSquare() const { return nRows == nColumns; }
    bool isNotEmpty() const { return nRows > 0 && nColumns > 0; }

    virtual std::unique_ptr<SparseMtrx> clone() const { OOFEM_ERROR("Not implemented"); }
    virtual void times(const FloatArray &x, FloatArray &answer) const { OOFEM_ERROR("Not implemented"); }
    virtual void timesT(const FloatArray &x, FloatArray &answer) const { OOFEM_ERROR("Not implemented"); }
    virtual void times(const FloatMatrix &B, FloatMatrix &answer) const { OOFEM_ERROR("Not implemented"); }
    virtual void times(double x) { OOFEM_ERROR("Not implemented"); }
    virtual void add(double x, SparseMtrx &m) { OOFEM_ERROR("Not implemented"); }
    virtual void addDiagonal(double x, FloatArray &m)
    {
        for (int i = 1; i <= m.giveSize(); ++i) this->at(i, i) += x * m.at(i);
    }
    virtual int buildInternalStructure(EngngModel *eModel, int di, const UnknownNumberingScheme &s) = 0;
    virtual int assemble(const IntArray &loc, const FloatMatrix &mat) = 0;
    virtual int assemble(const IntArray &rloc, const IntArray &cloc, const FloatMatrix &mat) = 0;
    virtual int assembleBegin() { return 1; }
    virtual int assembleEnd() { return 1; }
    virtual bool canBeFactorized() const = 0;
    virtual void zero() = 0;
    virtual double &at(int i, int j) = 0;
    virtual double at(int i, int j) const = 0;
    virtual bool isAllocatedAt(int i, int j) const { return false; }
    virtual void printStatistics() const {}
    virtual SparseMtrxType giveType() const = 0;
    virtual bool isAsymmetric() const = 0;
    virtual const char *giveClassName() const = 0;
};

class SparseLinearSystemNM {
protected:
    Domain *domain;
    EngngModel *engngModel;

public:
    SparseLinearSystemNM(Domain *d, EngngModel *m) : domain(d), engngModel(m) {}
    virtual ~SparseLinearSystemNM() {}
    virtual void initializeFrom(InputRecord &ir) {}
    virtual ConvergedReason solve(SparseMtrx &A, FloatArray &b, FloatArray &x) = 0;
    // several right-hand sides (columns of B), one at a time; X holds the
    // initial guesses (sparselinsystemnm.C:47-70)
    virtual ConvergedReason solve(SparseMtrx &A, FloatMatrix &B, FloatMatrix &X);
    virtual LinSystSolverType giveLinSystSolverType() const = 0;
    virtual SparseMtrxType giveRecommendedMatrix(bool symmetric) const { return SMT_Skyline; }
    virtual const char *giveClassName() const = 0;
};

}  // namespace oofem

// HipCSRSparseMtrx: the SparseMtrx of type SMT_HipCSR.  Storage and every
// product live on the GPU (libhipfem, include/hipfem.h); this class only maps
// the reference's SparseMtrx contract (sparsemtrx.h:60-281) onto the C ABI:
//   buildInternalStructure  pattern from the location arrays of the domain's
//                           elements (compcol.C:174-197: giveDomain,
//                           giveElements, giveLocationArray), or from the
//                           batched evaluator when the model provides one
//   assemble(loc, mat)      staged on the host, sent at assembleEnd()
//   times / timesT          device SpMV
//   at(i,j) const           single-entry read; at(i,j) non-const works on a
//                           host mirror that is written back before the next
//                           device use, and bumps the version counter
#pragma once

#include <vector>

#include "../../include/hipfem.h"
#include "engngm.h"
#include "sparsemtrx.h"

namespace oofem {

// one context per process (device from HIPFEM_DEVICE, default 0); every
// plug-in object created without an explicit context shares it
hipfem_ctx *hipfemDefaultContext();
void hipfemCheck(int rc, const char *what);

// Engineering models that evaluate their elements in batches implement this
// interface (no counterpart in the reference): the matrix then takes its
// pattern from the evaluator instead of one location array per element.
class BatchedEvaluatorProvider {
public:
    virtual ~BatchedEvaluatorProvider() {}
    virtual hfem_model *giveBatchedEvaluator(int di) = 0;
};

class HipCSRSparseMtrx : public SparseMtrx {
    hipfem_ctx *ctx;
    hipcsr_mtx *m = nullptr;
    mutable std::vector<double> mirror;   // host copy for non-const at()
    mutable bool mirrorValid = false;
    bool mirrorDirty = false;

    void flushMirror();
    void syncVersion();

public:
    HipCSRSparseMtrx(int n = 0, int mm = 0);
    explicit HipCSRSparseMtrx(hipfem_ctx *c);
    ~HipCSRSparseMtrx() override;

    std::unique_ptr<SparseMtrx> clone() const override;
    void times(const FloatArray &x, FloatArray &answer) const override;
    void timesT(const FloatArray &x, FloatArray &answer) const override;
    void times(const FloatMatrix &B, FloatMatrix &answer) const override;
    void times(double x) override;
    void add(double x, SparseMtrx &m) override;
    void addDiagonal(double x, FloatArray &d) override;
    int buildInternalStructure(EngngModel *eModel, int di, const UnknownNumberingScheme &s) override;
    int assemble(const IntArray &loc, const FloatMatrix &mat) override;
    int assemble(const IntArray &rloc, const IntArray &cloc, const FloatMatrix &mat) override;
    int assembleBegin() override;
    int assembleEnd() override;
    bool canBeFactorized() const override { return false; }
    void zero() override;
    double &at(int i, int j) override;
    double at(int i, int j) const override;
    bool isAllocatedAt(int i, int j) const override;
    void printStatistics() const override;
    SparseMtrxType giveType() const override { return SMT_HipCSR; }
    bool isAsymmetric() const override { return true; }   // full storage
    const char *giveClassName() const override { return "HipCSRSparseMtrx"; }

    // GPU-side access for the solver and the batched evaluator
    hipcsr_mtx *handle();
    hipfem_ctx *context() const { return ctx; }
    long deviceVersion() const { return hipcsr_version(m); }
};

}  // namespace oofem

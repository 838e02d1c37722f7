// Enum-keyed registry for sparse matrices and linear solvers, mirroring the
// reference's ClassFactory entries for these two kinds
// (src/core/classfactory.h:149-150 REGISTER_SparseMtrx / REGISTER_SparseLinSolver,
// classfactory.C:145-170 create/register).  Decks select them by integer
// (smtype / lstype), so the appended SMT_HipCSR = 11 and ST_HipPCG = 9 are
// what an unchanged deck switches to.
#pragma once

#include <functional>
#include <map>
#include <memory>

#include "sparsemtrx.h"

namespace oofem {

class ClassFactory {
    std::map<int, std::function<std::unique_ptr<SparseMtrx>()>> sparseMtrxList;
    std::map<int, std::function<std::unique_ptr<SparseLinearSystemNM>(Domain *, EngngModel *)>> sparseLinSolList;

public:
    bool registerSparseMtrx(SparseMtrxType type, std::function<std::unique_ptr<SparseMtrx>()> f)
    {
        sparseMtrxList[(int)type] = f;
        return true;
    }
    bool registerSparseLinSolver(LinSystSolverType type,
                                 std::function<std::unique_ptr<SparseLinearSystemNM>(Domain *, EngngModel *)> f)
    {
        sparseLinSolList[(int)type] = f;
        return true;
    }
    std::unique_ptr<SparseMtrx> createSparseMtrx(SparseMtrxType type)
    {
        auto it = sparseMtrxList.find((int)type);
        return it == sparseMtrxList.end() ? nullptr : it->second();
    }
    std::unique_ptr<SparseLinearSystemNM> createSparseLinSolver(LinSystSolverType st, Domain *d, EngngModel *m)
    {
        auto it = sparseLinSolList.find((int)st);
        return it == sparseLinSolList.end() ? nullptr : it->second(d, m);
    }
};

ClassFactory &GiveClassFactory();

#define REGISTER_SparseMtrx(class, type)                                                         \
    static bool __dummy_##class __attribute__((unused)) = GiveClassFactory().registerSparseMtrx( \
        type, []() -> std::unique_ptr<SparseMtrx> { return std::make_unique<class>(); });
#define REGISTER_SparseLinSolver(class, type)                                                          \
    static bool __dummy_##class __attribute__((unused)) = GiveClassFactory().registerSparseLinSolver( \
        type, [](Domain *d, EngngModel *m) -> std::unique_ptr<SparseLinearSystemNM> {                 \
            return std::make_unique<class>(d, m);                                                      \
        });

}  // namespace oofem

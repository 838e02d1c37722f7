// StructuralEngngModel (src/sm/EngineeringModels/structengngmodel.h:167):
// the structural models' override of updateInternalRHS, the internal-force
// hook of NRSolver's InternalRhs component (structengngmodel.C:215-233).
#pragma once

#include "engngm.h"

namespace oofem {

class StructuralEngngModel : public EngngModel {
public:
    void updateInternalRHS(FloatArray &answer, TimeStep *tStep, Domain *d, FloatArray *eNorm) override;
};

// the reference assembles InternalForceAssembler vectors element by element
// here; the mirror has no element library, so a model must override it
inline void StructuralEngngModel::updateInternalRHS(FloatArray &answer, TimeStep *tStep, Domain *d, FloatArray *eNorm)
{
    OOFEM_ERROR("no element library in this build: the engineering model must evaluate its internal forces");
}

}  // namespace oofem

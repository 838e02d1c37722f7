// Keywords of the IML solver (src/core/iml/imlsolver.h:45-52), the ones a
// deck gives to lstype 1 and that hippcg reads with the same meaning.
#pragma once

#define _IFT_IMLSolver_Name "iml"
#define _IFT_IMLSolver_stype "stype"
#define _IFT_IMLSolver_lstol "lstol"
#define _IFT_IMLSolver_lsiter "lsiter"
#define _IFT_IMLSolver_lsprecond "lsprecond"

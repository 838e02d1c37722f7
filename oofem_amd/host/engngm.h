// The engineering-model side of the plug-in boundary, restated with the
// reference's names and signatures (only what the plug-ins and the drivers
// here touch):
//   TimeStep          src/core/timestep.h:81-213 (number, target time,
//                     solution state counter)
//   Element           src/core/element.h:228 giveLocationArray
//   Domain            src/core/domain.h:293-294, 463 element access
//   MatResponseMode   src/core/matresponsemode.h:41-45 (the stiffness kinds)
//   MatrixAssembler / TangentAssembler
//                     src/core/assemblercallback.h:84-95, 185-197
//   EngngModel        src/core/engngm.h:356 giveDomain, :527
//                     giveNumberOfDomainEquations, :852 updateInternalRHS,
//                     :957-958 assemble -- the two virtual hooks a batched
//                     evaluator overrides
// The mirror has no element library: an Element answers its location array
// through its Domain, and MatrixAssembler::matrixFromElement, which would
// call the element's characteristic matrix, raises an error.
#pragma once

#include <functional>
#include <memory>
#include <vector>

#include "sparsemtrx.h"

namespace oofem {

class TimeStep {
public:
    typedef unsigned long StateCounterType;

protected:
    int number;
    double targetTime;
    StateCounterType solutionStateCounter = 0;

public:
    TimeStep(int n = 0, double t = 0.0) : number(n), targetTime(t) {}
    int giveNumber() { return number; }
    double giveTargetTime() { return targetTime; }
    void setNumber(int n) { number = n; }
    void setTargetTime(double t) { targetTime = t; }
    StateCounterType giveSolutionStateCounter() { return solutionStateCounter; }
    void incrementStateCounter() { solutionStateCounter++; }
};

class Domain;

class Element {
    int number;
    Domain *domain;

public:
    Element(int n, Domain *d) : number(n), domain(d) {}
    virtual ~Element() {}
    int giveNumber() const { return number; }
    Domain *giveDomain() const { return domain; }
    void giveLocationArray(IntArray &locationArray, const UnknownNumberingScheme &s, IntArray *dofIds = NULL) const;
};

class Domain {
    int number;
    std :: vector< std :: unique_ptr< Element > > elementList;
    // location array of element n (1-based) in the default numbering, and
    // the dof id (1 u, 2 v, 3 w) of each entry
    std::function<void(int, IntArray &, IntArray *)> locator;

public:
    explicit Domain(int n = 1) : number(n) {}
    int giveNumber() const { return number; }
    // the engineering model describes its elements by count and locator
    void setElements(int nelem, std::function<void(int, IntArray &, IntArray *)> f)
    {
        locator = std::move(f);
        elementList.clear();
        elementList.reserve((size_t)nelem);
        for (int i = 1; i <= nelem; i++) elementList.push_back(std::make_unique<Element>(i, this));
    }
    Element *giveElement(int n) { return elementList[(size_t)n - 1].get(); }
    std :: vector< std :: unique_ptr< Element > > &giveElements() { return this->elementList; }
    int giveNumberOfElements() const { return (int)elementList.size(); }
    void elementLocation(int n, IntArray &loc, IntArray *dofIds) const
    {
        if (!locator) OOFEM_ERROR("domain %d describes no elements", number);
        locator(n, loc, dofIds);
    }
};

inline void Element::giveLocationArray(IntArray &locationArray, const UnknownNumberingScheme &s, IntArray *dofIds) const
{
    domain->elementLocation(number, locationArray, dofIds);
}

enum MatResponseMode { TangentStiffness = 0, SecantStiffness = 1, ElasticStiffness = 2 };

class MatrixAssembler {
public:
    virtual ~MatrixAssembler() {}
    virtual void matrixFromElement(FloatMatrix &mat, Element &element, TimeStep *tStep) const
    {
        OOFEM_ERROR("no element library in this build: element %d has no characteristic matrix", element.giveNumber());
    }
    virtual void locationFromElement(IntArray &loc, Element &element, const UnknownNumberingScheme &s, IntArray *dofIds = nullptr) const
    {
        element.giveLocationArray(loc, s, dofIds);
    }
};

class TangentAssembler : public MatrixAssembler {
protected:
    MatResponseMode rmode;

public:
    TangentAssembler(MatResponseMode m = TangentStiffness): MatrixAssembler(), rmode(m) {}
};

class EngngModel {
protected:
    std :: vector< std :: unique_ptr< Domain > > domainList;

public:
    EngngModel() {}
    virtual ~EngngModel() {}

    Domain *giveDomain(int n);
    virtual int giveNumberOfDomainEquations(int di, const UnknownNumberingScheme &num) = 0;

    virtual void updateInternalRHS(FloatArray &answer, TimeStep *tStep, Domain *d, FloatArray *eNorm);

    virtual void assemble(SparseMtrx &answer, TimeStep *tStep,
                          const MatrixAssembler &ma, const UnknownNumberingScheme &s, Domain *domain);
};

inline Domain *EngngModel::giveDomain(int n)
{
    if (n < 1 || n > (int)domainList.size()) OOFEM_ERROR("undefined domain %d", n);
    return domainList[(size_t)n - 1].get();
}

// the element loop of EngngModel::assemble (engngm.C:889-1037): one
// characteristic matrix per element, scattered through its location array
inline void EngngModel::assemble(SparseMtrx &answer, TimeStep *tStep, const MatrixAssembler &ma,
                                 const UnknownNumberingScheme &s, Domain *domain)
{
    IntArray loc;
    FloatMatrix mat;
    answer.assembleBegin();
    for (auto &elem : domain->giveElements()) {
        ma.matrixFromElement(mat, *elem, tStep);
        if (mat.isNotEmpty()) {
            ma.locationFromElement(loc, *elem, s);
            if (answer.assemble(loc, mat) == 0) OOFEM_ERROR("sparse matrix assemble error");
        }
    }
    answer.assembleEnd();
}

inline void EngngModel::updateInternalRHS(FloatArray &answer, TimeStep *tStep, Domain *d, FloatArray *eNorm)
{
    OOFEM_ERROR("Unknown Type of component.");
}

}  // namespace oofem

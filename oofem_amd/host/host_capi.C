// C entry points of the host layer (libhipfem_host.so), used by the Python
// driver and tests: a StructuralProblem created from flat arrays, solved
// through the class factory (smtype / lstype) and the mirrored drivers.
#include <algorithm>
#include <cstring>
#include <string>

#include "../../include/hipfem_host.h"
#include "structmodel.h"

using namespace oofem;

namespace {
thread_local std::string g_host_err;

template <typename F>
int guarded(F &&f)
{
    try {
        f();
        return 0;
    } catch (const std::exception &e) {
        g_host_err = e.what();
        return -1;
    }
}
}  // namespace

extern "C" {

const char *hfem_problem_last_error(void) { return g_host_err.c_str(); }

int hfem_problem_create(int nnode, const double *coords, int etype, int nelem, const int *conn, const int *emat,
                        int nmat, const double *mat, const int *eqn, int neq, int npeq, hfem_problem **out)
{
    return guarded([&] {
        *out = reinterpret_cast<hfem_problem *>(
            new StructuralProblem(nnode, coords, etype, nelem, conn, emat, nmat, mat, eqn, neq, npeq));
    });
}

void hfem_problem_destroy(hfem_problem *p) { delete reinterpret_cast<StructuralProblem *>(p); }

// smtype / lstype: deck integers (SMT_HipCSR = 11, ST_HipPCG = 9); the
// remaining fields are the IMLSolver / hippcg keywords and NRSolver options
int hfem_problem_set_options(hfem_problem *p, int smtype, int lstype, double lstol, int lsiter, int lsprecond,
                             int mixed, double irtol, double rtolf, int maxiter, int manrmsteps)
{
    return guarded([&] {
        auto *P = reinterpret_cast<StructuralProblem *>(p);
        P->setSparseMtrxType(smtype);
        P->setLinSystSolverType(lstype);
        DynamicInputRecord &ir = P->giveSolverRecord();
        ir.setField(lstol, _IFT_IMLSolver_lstol);
        ir.setField(lsiter, _IFT_IMLSolver_lsiter);
        ir.setField(lsprecond, _IFT_IMLSolver_lsprecond);
        ir.setField(mixed, _IFT_HipPCGSolver_mixed);
        ir.setField(irtol, _IFT_HipPCGSolver_irtol);
        P->nr.rtolf = rtolf;
        if (maxiter > 0) P->nr.maxiter = maxiter;
        P->nr.manrmsteps = manrmsteps;
    });
}

int hfem_problem_set_forcing(hfem_problem *p, int on, double eta_max)
{
    return guarded([&] {
        if (on && !(eta_max > 0.0 && eta_max < 1.0)) OOFEM_ERROR("forcing term bound must be in (0, 1), got %g", eta_max);
        auto *P = reinterpret_cast<StructuralProblem *>(p);
        P->nr.forcing = on != 0;
        if (on) P->nr.forcingMax = eta_max;
    });
}

int hfem_problem_set_initial_guess(hfem_problem *p, int type)
{
    return guarded([&] {
        if (type != IG_None && type != IG_Tangent) OOFEM_ERROR("Initial guess type: %d not supported", type);
        reinterpret_cast<StructuralProblem *>(p)->initialGuessType = (InitialGuess)type;
    });
}

int hfem_problem_set_prescribed(hfem_problem *p, const double *up)
{
    return guarded([&] { reinterpret_cast<StructuralProblem *>(p)->setPrescribedValues(up); });
}

int hfem_problem_set_load(hfem_problem *p, const double *f_free, const double *f_presc)
{
    return guarded([&] { reinterpret_cast<StructuralProblem *>(p)->setExternalLoad(f_free, f_presc); });
}

int hfem_problem_linear_static(hfem_problem *p, int *iterations, double *residual)
{
    return guarded([&] { reinterpret_cast<StructuralProblem *>(p)->solveLinearStatic(*iterations, *residual); });
}

// out[0] Newton iterations, out[1] PCG iterations, out[2] status; err = final force error
int hfem_problem_nonlinear_step(hfem_problem *p, int *out3, double *force_error)
{
    return guarded([&] {
        NRStepReport r = reinterpret_cast<StructuralProblem *>(p)->solveNonlinearStep();
        out3[0] = r.iterations;
        out3[1] = r.linearIterations;
        out3[2] = (int)r.status;
        *force_error = r.forceError;
    });
}

int hfem_problem_set_arc_length(hfem_problem *p, double psi, int maxiter, int maxrestarts, double minsteplength,
                                double steplength, double initialsteplength, int reqiterations, int miniterations,
                                int manrmsteps, double rtolf, double rtold, int rootselectiontype)
{
    return guarded([&] {
        CALMOptions o;
        o.psi = psi;
        o.maxiter = maxiter;
        o.maxRestarts = maxrestarts;
        o.minStepLength = minsteplength;
        o.maxStepLength = steplength;
        o.initialStepLength = initialsteplength;
        o.reqIterations = reqiterations;
        o.minIterations = miniterations;
        o.manrmsteps = manrmsteps;
        o.rtolf = rtolf;
        o.rtold = rtold;
        o.rootSelection = rootselectiontype;
        reinterpret_cast<StructuralProblem *>(p)->setArcLength(o);
    });
}

int hfem_problem_set_arc_length_hpc(hfem_problem *p, int hpcmode, int n, const int *eqs, const double *weights)
{
    return guarded([&] {
        if (n < 0 || (n > 0 && !eqs) || (hpcmode == 2 && n > 0 && !weights)) OOFEM_ERROR("hpc: bad arguments");
        std::vector<int> e(eqs, eqs + n);
        std::vector<double> w;
        if (hpcmode == 2) w.assign(weights, weights + n);
        reinterpret_cast<StructuralProblem *>(p)->setArcLengthHpc(hpcmode, e, w);
    });
}

int hfem_problem_set_reference_load(hfem_problem *p, const double *f_free, const double *f_presc)
{
    return guarded([&] {
        auto *P = reinterpret_cast<StructuralProblem *>(p);
        if (!f_free || (P->giveNpeq() > 0 && !f_presc)) OOFEM_ERROR("reference load: missing vector");
        P->setReferenceLoad(f_free, f_presc);
    });
}

int hfem_problem_arc_length_step(hfem_problem *p, int *out4, double *out4d)
{
    return guarded([&] {
        CALMStepReport r = reinterpret_cast<StructuralProblem *>(p)->solveArcLengthStep();
        out4[0] = r.iterations;
        out4[1] = r.linearIterations;
        out4[2] = r.restarts;
        out4[3] = (int)r.status;
        out4d[0] = r.lambda;
        out4d[1] = r.stepLength;
        out4d[2] = r.forceError;
        out4d[3] = r.dispError;
    });
}

int hfem_problem_set_nr_extras(hfem_problem *p, int linesearch, double ls_tol, double ls_amp, double ls_maxeta,
                               double maxinc)
{
    return guarded([&] {
        if (!(maxinc > 0.0)) OOFEM_ERROR("maxinc must be positive, got %g", maxinc);
        auto &nr = reinterpret_cast<StructuralProblem *>(p)->nr;
        nr.lineSearch = linesearch != 0;
        nr.lsTol = ls_tol;
        nr.lsAmp = ls_amp;
        nr.lsMaxEta = ls_maxeta;
        nr.maxInc = maxinc;
    });
}

int hfem_problem_set_nr_criteria(hfem_problem *p, double rtold, int miniter, int ls_keep_point)
{
    return guarded([&] {
        if (miniter < 0) OOFEM_ERROR("miniter must be >= 0, got %d", miniter);
        auto &nr = reinterpret_cast<StructuralProblem *>(p)->nr;
        nr.rtold = rtold;
        nr.minIterations = miniter;
        nr.lsKeepPoint = ls_keep_point != 0;
    });
}

int hfem_problem_nr_errors(hfem_problem *p, double *out, int cap, int *n)
{
    return guarded([&] {
        const std::vector<double> &e = reinterpret_cast<StructuralProblem *>(p)->giveLastNRErrors();
        *n = (int)(e.size() / 6);
        for (int i = 0; i < std::min(cap, *n) * 6; i++) out[i] = e[(size_t)i];
    });
}

int hfem_problem_update_internal_rhs(hfem_problem *p, double *f, double *enorm3)
{
    return guarded([&] {
        auto *P = reinterpret_cast<StructuralProblem *>(p);
        EngngModel *em = P;   // through the reference's virtual hook
        FloatArray answer, eNorm;
        em->updateInternalRHS(answer, P->giveCurrentStep(), em->giveDomain(1), &eNorm);
        std::memcpy(f, answer.givePointer(), sizeof(double) * (size_t)answer.giveSize());
        for (int k = 0; k < 3; k++) enorm3[k] = eNorm.at(k + 1);
    });
}

int hfem_problem_assemble(hfem_problem *p, int response_mode)
{
    return guarded([&] {
        auto *P = reinterpret_cast<StructuralProblem *>(p);
        if (response_mode < 0 || response_mode > 2) OOFEM_ERROR("MatResponseMode %d is not a stiffness", response_mode);
        SparseMtrx *K = P->giveStiffnessMatrix();
        if (!K) OOFEM_ERROR("the problem has no stiffness matrix yet");
        EngngModel *em = P;
        em->assemble(*K, P->giveCurrentStep(), TangentAssembler((MatResponseMode)response_mode),
                     EModelDefaultEquationNumbering(), em->giveDomain(1));
    });
}

int hfem_problem_set_ddm(hfem_problem *p, int n, const int *eqs, const double *values)
{
    return guarded([&] {
        auto *P = reinterpret_cast<StructuralProblem *>(p);
        if (n < 0 || (n > 0 && (!eqs || !values))) OOFEM_ERROR("ddm: bad arguments");
        for (int k = 0; k < n; k++)
            if (eqs[k] < 1 || eqs[k] > P->giveNeq()) OOFEM_ERROR("ddm: equation %d is not a free equation", eqs[k]);
        P->nr.ddmEq.assign(eqs, eqs + n);
        P->nr.ddmVal.assign(values, values + n);
    });
}

int hfem_problem_save_converged(hfem_problem *p)
{
    return guarded([&] { reinterpret_cast<StructuralProblem *>(p)->saveConverged(); });
}

int hfem_problem_restart_step(hfem_problem *p)
{
    return guarded([&] { reinterpret_cast<StructuralProblem *>(p)->restartStep(); });
}

int hfem_problem_set_constrained_nr(hfem_problem *p, int miniter)
{
    return guarded([&] { reinterpret_cast<StructuralProblem *>(p)->nr.constrainedMinIter = miniter; });
}

int hfem_problem_set_ddm_factor(hfem_problem *p, double factor)
{
    return guarded([&] { reinterpret_cast<StructuralProblem *>(p)->nr.ddmFactor = factor; });
}

int hfem_problem_set_stiffness_mode(hfem_problem *p, int mode)
{
    return guarded([&] {
        if (mode < 0 || mode > 3) OOFEM_ERROR("unknown stiffness mode %d", mode);
        reinterpret_cast<StructuralProblem *>(p)->nr.stiffMode = mode;
    });
}

int hfem_problem_run_steps(hfem_problem *p, const double *up_ref, const double *f_free, const double *f_presc,
                           int nltf, const double *ltf_t, const double *ltf_v, double final_time, double dt,
                           int reduction, double dtmin, double dtmax, double ncrf, int nmaxrestarts,
                           int min_required_iter, int max_required_iter, int max_records, double *rec_time,
                           double *rec_dt, int *rec_int, int *nrec)
{
    return guarded([&] {
        auto *P = reinterpret_cast<StructuralProblem *>(p);
        if (nltf < 0 || (nltf > 0 && (!ltf_t || !ltf_v))) OOFEM_ERROR("load time function: bad arguments");
        if (P->giveNpeq() > 0 && !up_ref) OOFEM_ERROR("reference prescribed values missing");
        ReductionStrategy rs;
        rs.on = reduction != 0;
        if (rs.on) {
            if (!(dtmin >= 0.0) || !(dtmax >= dtmin) || !(ncrf > 0.0 && ncrf < 1.0) || nmaxrestarts < 0)
                OOFEM_ERROR("reduction strategy: bad values (dtmin %g, dtmax %g, ncrf %g, nmaxrestarts %d)", dtmin,
                            dtmax, ncrf, nmaxrestarts);
            rs.dtmin = dtmin;
            rs.dtmax = dtmax;
            rs.ncrf = ncrf;
            rs.nmaxrestarts = nmaxrestarts;
            rs.minRequiredIter = min_required_iter;
            rs.maxRequiredIter = max_required_iter;
        }
        std::vector<double> tt(ltf_t, ltf_t + nltf), vv(ltf_v, ltf_v + nltf);
        std::vector<double> zero((size_t)std::max(1, P->giveNpeq()), 0.0);
        auto recs = P->solveTimeSteps(up_ref ? up_ref : zero.data(), f_free, f_presc, tt, vv, final_time, dt, rs);
        const int n = (int)std::min<size_t>(recs.size(), (size_t)std::max(0, max_records));
        for (int i = 0; i < n; i++) {
            if (rec_time) rec_time[i] = recs[i].time;
            if (rec_dt) rec_dt[i] = recs[i].dt;
            if (rec_int) {
                int *r = rec_int + 6 * i;
                r[0] = recs[i].newton;
                r[1] = recs[i].pcg;
                r[2] = recs[i].status;
                r[3] = recs[i].restarts;
                r[4] = recs[i].lineSearches;
                r[5] = 0;
            }
        }
        *nrec = (int)recs.size();
    });
}

int hfem_problem_update(hfem_problem *p)
{
    return guarded([&] { reinterpret_cast<StructuralProblem *>(p)->updateYourself(); });
}

int hfem_problem_get_solution(hfem_problem *p, double *u)
{
    return guarded([&] { reinterpret_cast<StructuralProblem *>(p)->giveSolution(u); });
}

int hfem_problem_get_reactions(hfem_problem *p, double *r)
{
    return guarded([&] { reinterpret_cast<StructuralProblem *>(p)->giveReactions(r); });
}

int hfem_problem_get_state(hfem_problem *p, double *s)
{
    return guarded([&] { reinterpret_cast<StructuralProblem *>(p)->giveGaussPointState(s); });
}

// SparseMtrx access through the factory-made matrix (at() / times())
int hfem_problem_matrix_at(hfem_problem *p, int i, int j, double *v)
{
    return guarded([&] {
        SparseMtrx *K = reinterpret_cast<StructuralProblem *>(p)->giveStiffnessMatrix();
        if (!K) OOFEM_ERROR("no stiffness matrix yet");
        *v = static_cast<const SparseMtrx *>(K)->at(i, j);
    });
}

int hfem_problem_set_profiling(hfem_problem *p, int on)
{
    return guarded([&] { reinterpret_cast<StructuralProblem *>(p)->setProfiling(on != 0); });
}

int hfem_problem_profile(hfem_problem *p, int *n64, double *ms64, int *n32, double *ms32)
{
    return guarded([&] { reinterpret_cast<StructuralProblem *>(p)->giveProfile(n64, ms64, n32, ms32); });
}

int hfem_problem_solver_info(hfem_problem *p, int *outer, int *matrix_free)
{
    return guarded([&] {
        auto *S = dynamic_cast<HipPCGSolver *>(reinterpret_cast<StructuralProblem *>(p)->giveLinearSolver());
        *outer = S ? hippcg_last_outer(S->handle()) : 0;
        *matrix_free = S ? hippcg_last_matrix_free(S->handle()) : 0;
    });
}

int hfem_problem_context(hfem_problem *p, hipfem_ctx **ctx)
{
    return guarded([&] { *ctx = reinterpret_cast<StructuralProblem *>(p)->giveContext(); });
}

int hfem_problem_set_surrogate(hfem_problem *p, hfem_mlp *net, double band, double tol)
{
    return guarded([&] {
        hipfemCheck(hfem_model_set_surrogate(reinterpret_cast<StructuralProblem *>(p)->giveBatchedEvaluator(1), net,
                                             band, tol),
                    "set_surrogate");
    });
}

int hfem_problem_surrogate_stats(hfem_problem *p, long long *counts3, int reset)
{
    return guarded([&] {
        int64_t c[3];
        hipfemCheck(hfem_model_surrogate_stats(reinterpret_cast<StructuralProblem *>(p)->giveBatchedEvaluator(1), c,
                                               reset),
                    "surrogate_stats");
        for (int i = 0; i < 3; i++) counts3[i] = c[i];
    });
}

int hfem_problem_matrix_layout(hfem_problem *p, int *form, long long *groups, long long *blocks)
{
    return guarded([&] {
        auto *H = dynamic_cast<HipCSRSparseMtrx *>(reinterpret_cast<StructuralProblem *>(p)->giveStiffnessMatrix());
        if (!H) OOFEM_ERROR("no hipcsr stiffness matrix yet");
        int64_t g = 0, b = 0;
        hipfemCheck(hipcsr_nodeblock_sizes(H->handle(), &g, &b), "nodeblock_sizes");
        *form = hipcsr_is_nodeblocked(H->handle());
        *groups = g;
        *blocks = b;
    });
}

int hfem_problem_matrix_nnz(hfem_problem *p, long long *nnz)
{
    return guarded([&] {
        auto *H = dynamic_cast<HipCSRSparseMtrx *>(reinterpret_cast<StructuralProblem *>(p)->giveStiffnessMatrix());
        *nnz = H ? (long long)hipcsr_nnz(H->handle()) : 0;
    });
}

int hfem_problem_matrix_class(hfem_problem *p, char *buf, int len)
{
    return guarded([&] {
        SparseMtrx *K = reinterpret_cast<StructuralProblem *>(p)->giveStiffnessMatrix();
        const char *n = K ? K->giveClassName() : "";
        std::strncpy(buf, n, (size_t)len - 1);
        buf[len - 1] = 0;
    });
}

}  // extern "C"

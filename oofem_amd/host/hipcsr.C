#include "hipcsr.h"

#include <algorithm>
#include <cstdlib>
#include <mutex>

#include "classfactory.h"

namespace oofem {

REGISTER_SparseMtrx(HipCSRSparseMtrx, SMT_HipCSR)

void hipfemCheck(int rc, const char *what)
{
    if (rc != 0) OOFEM_ERROR("%s failed (status %d): %s", what, rc, hipfem_last_error());
}

hipfem_ctx *hipfemDefaultContext()
{
    static hipfem_ctx *ctx = nullptr;
    static std::once_flag once;
    std::call_once(once, [] {
        const char *dev = std::getenv("HIPFEM_DEVICE");
        hipfemCheck(hipfem_ctx_create(dev ? std::atoi(dev) : 0, &ctx), "hipfem_ctx_create");
    });
    return ctx;
}

HipCSRSparseMtrx::HipCSRSparseMtrx(int n, int mm) : SparseMtrx(n, mm), ctx(hipfemDefaultContext())
{
    hipfemCheck(hipcsr_create(ctx, &m), "hipcsr_create");
}

HipCSRSparseMtrx::HipCSRSparseMtrx(hipfem_ctx *c) : SparseMtrx(0, 0), ctx(c)
{
    hipfemCheck(hipcsr_create(ctx, &m), "hipcsr_create");
}

HipCSRSparseMtrx::~HipCSRSparseMtrx()
{
    if (m) hipcsr_destroy(m);
}

void HipCSRSparseMtrx::syncVersion() { this->version = hipcsr_version(m); }

void HipCSRSparseMtrx::flushMirror()
{
    if (mirrorDirty) {
        hipfemCheck(hipcsr_upload_values(m, mirror.data()), "upload");
        mirrorDirty = false;
        syncVersion();
    }
}

hipcsr_mtx *HipCSRSparseMtrx::handle()
{
    flushMirror();
    mirrorValid = false;  // the device copy may change under the caller
    return m;
}

std::unique_ptr<SparseMtrx> HipCSRSparseMtrx::clone() const
{
    const_cast<HipCSRSparseMtrx *>(this)->flushMirror();
    auto c = std::make_unique<HipCSRSparseMtrx>(ctx);
    hipcsr_destroy(c->m);
    c->m = nullptr;
    hipfemCheck(hipcsr_clone(m, &c->m), "clone");
    c->nRows = nRows;
    c->nColumns = nColumns;
    c->version = version;
    return c;
}

void HipCSRSparseMtrx::times(const FloatArray &x, FloatArray &answer) const
{
    const_cast<HipCSRSparseMtrx *>(this)->flushMirror();
    if (x.giveSize() != nColumns) OOFEM_ERROR("incompatible dimensions");
    answer.resize(nRows);
    hipfemCheck(hipcsr_times(m, x.givePointer(), answer.givePointer()), "times");
}

void HipCSRSparseMtrx::timesT(const FloatArray &x, FloatArray &answer) const
{
    const_cast<HipCSRSparseMtrx *>(this)->flushMirror();
    if (x.giveSize() != nRows) OOFEM_ERROR("incompatible dimensions");
    answer.resize(nColumns);
    hipfemCheck(hipcsr_times_t(m, x.givePointer(), answer.givePointer()), "timesT");
}

void HipCSRSparseMtrx::times(const FloatMatrix &B, FloatMatrix &answer) const
{
    const int nc = B.giveNumberOfColumns();
    answer.resize(nRows, nc);
    FloatArray x(nColumns), y;
    for (int c = 1; c <= nc; c++) {
        for (int i = 1; i <= nColumns; i++) x.at(i) = B.at(i, c);
        times(x, y);
        for (int i = 1; i <= nRows; i++) answer.at(i, c) = y.at(i);
    }
}

void HipCSRSparseMtrx::times(double x)
{
    flushMirror();
    mirrorValid = false;
    hipfemCheck(hipcsr_scale(m, x), "times(double)");
    syncVersion();
}

void HipCSRSparseMtrx::add(double x, SparseMtrx &other)
{
    auto *o = dynamic_cast<HipCSRSparseMtrx *>(&other);
    if (!o) OOFEM_ERROR("add: operand is not a HipCSRSparseMtrx");
    flushMirror();
    o->flushMirror();
    mirrorValid = false;
    hipfemCheck(hipcsr_add(m, x, o->m), "add");
    syncVersion();
}

void HipCSRSparseMtrx::addDiagonal(double x, FloatArray &d)
{
    flushMirror();
    mirrorValid = false;
    hipfemCheck(hipcsr_add_diagonal(m, x, d.givePointer()), "addDiagonal");
    syncVersion();
}

int HipCSRSparseMtrx::buildInternalStructure(EngngModel *eModel, int di, const UnknownNumberingScheme &s)
{
    const int neq = eModel->giveNumberOfDomainEquations(di, s);
    mirror.clear();
    mirrorValid = mirrorDirty = false;
    auto *bp = dynamic_cast<BatchedEvaluatorProvider *>(eModel);
    if (hfem_model *ev = bp ? bp->giveBatchedEvaluator(di) : nullptr) {
        hipfemCheck(hfem_model_build_structure(ev, m), "buildInternalStructure");
    } else {
        // union of loc x loc over the elements (compcol.C:174-197); location
        // arrays of different lengths are padded with 0 (no equation)
        IntArray loc;
        Domain *domain = eModel->giveDomain(di);
        std::vector<IntArray> locs;
        locs.reserve((size_t)domain->giveNumberOfElements());
        int ndofe = 0;
        for (auto &elem : domain->giveElements()) {
            elem->giveLocationArray(loc, s);
            ndofe = std::max(ndofe, loc.giveSize());
            locs.push_back(loc);
        }
        std::vector<int> all(locs.size() * (size_t)ndofe, 0);
        for (size_t e = 0; e < locs.size(); e++)
            for (int i = 0; i < locs[e].giveSize(); i++) all[e * (size_t)ndofe + i] = locs[e][i];
        hipfemCheck(hipcsr_build_structure(m, neq, (int)locs.size(), ndofe, all.data()), "buildInternalStructure");
    }
    nRows = nColumns = neq;
    syncVersion();
    return 1;
}

int HipCSRSparseMtrx::assemble(const IntArray &loc, const FloatMatrix &mat)
{
    return assemble(loc, loc, mat);
}

int HipCSRSparseMtrx::assemble(const IntArray &rloc, const IntArray &cloc, const FloatMatrix &mat)
{
    flushMirror();
    mirrorValid = false;
    const int nr = rloc.giveSize(), nc = cloc.giveSize();
    if (mat.giveNumberOfRows() != nr || mat.giveNumberOfColumns() != nc) return 0;
    std::vector<double> rm((size_t)nr * nc);   // column-major FloatMatrix -> row-major
    for (int i = 1; i <= nr; i++)
        for (int j = 1; j <= nc; j++) rm[(size_t)(i - 1) * nc + (j - 1)] = mat.at(i, j);
    return hipcsr_assemble_rc(m, nr, rloc.givePointer(), nc, cloc.givePointer(), rm.data()) == 0 ? 1 : 0;
}

int HipCSRSparseMtrx::assembleBegin() { return hipcsr_assemble_begin(m) == 0 ? 1 : 0; }

int HipCSRSparseMtrx::assembleEnd()
{
    const int rc = hipcsr_assemble_end(m);
    syncVersion();
    return rc == 0 ? 1 : 0;
}

void HipCSRSparseMtrx::zero()
{
    mirrorDirty = false;
    mirrorValid = false;
    hipfemCheck(hipcsr_zero(m), "zero");
    syncVersion();
}

double &HipCSRSparseMtrx::at(int i, int j)
{
    checkBounds(i, j);
    int64_t pos = -1;
    hipfemCheck(hipcsr_find(m, i, j, &pos), "at");
    if (pos < 0) OOFEM_ERROR("at(%d,%d): entry not in the sparse structure", i, j);
    if (!mirrorValid) {
        // one download of the values per matrix version; later entries are
        // found in the host pattern in O(log row length)
        mirror.resize((size_t)hipcsr_nnz(m));
        hipfemCheck(hipcsr_download(m, nullptr, nullptr, mirror.data()), "download");
        mirrorValid = true;
    }
    mirrorDirty = true;
    this->version++;
    return mirror[(size_t)pos];
}

double HipCSRSparseMtrx::at(int i, int j) const
{
    checkBounds(i, j);
    const_cast<HipCSRSparseMtrx *>(this)->flushMirror();
    double v = 0.0;
    hipfemCheck(hipcsr_get(m, i, j, &v), "at");
    return v;
}

bool HipCSRSparseMtrx::isAllocatedAt(int i, int j) const { return hipcsr_is_allocated(m, i, j) != 0; }

void HipCSRSparseMtrx::printStatistics() const
{
    printf("HipCSRSparseMtrx: %d x %d, nnz %lld, version %ld\n", nRows, nColumns, (long long)hipcsr_nnz(m),
           (long)version);
}

}  // namespace oofem

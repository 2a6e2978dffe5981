// Mixed-precision tetrapyd Gram (BF16 hi+lo split operands) -- placeholder
// until the kernel lands; the FP64 path is complete.
#include "mcmb_internal.h"
namespace mcmb {
int gram_mixed_launch(mcmb_ctx* c, const long* d_rowptr, const double* d_ws, double* d_gamma)
{
    (void)c; (void)d_rowptr; (void)d_ws; (void)d_gamma;
    return fail(MCMB_ERR_STATE, "mixed-precision Gram not available in this build");
}
int gram_gl_launch(mcmb_ctx* c, int npt, double* d_gamma)
{
    (void)c; (void)npt; (void)d_gamma;
    return fail(MCMB_ERR_STATE, "GL-factorised Gram not available in this build");
}
}  // namespace mcmb
extern "C" int mcmb_gamma_gl(mcmb_ctx* c, int npt, double* gh, double* gd)
{
    (void)npt; (void)gh; (void)gd;
    if (!c) return mcmb::fail(MCMB_ERR_ARG, "null context");
    return mcmb::gram_gl_launch(c, npt, gd);
}
extern "C" int mcmb_orthonormalise(mcmb_ctx* c, int n, const double* g, double* o, double* l, double* nm, int* bad)
{
    (void)c; (void)n; (void)g; (void)o; (void)l; (void)nm; (void)bad;
    return mcmb::fail(MCMB_ERR_STATE, "orthonormalisation not available in this build");
}
extern "C" int mcmb_cholesky_solve(mcmb_ctx* c, int n, const double* f, int nrhs, const double* a, double* b)
{
    (void)c; (void)n; (void)f; (void)nrhs; (void)a; (void)b;
    return mcmb::fail(MCMB_ERR_STATE, "cholesky solve not available in this build");
}

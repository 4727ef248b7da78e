// Checker-only shim (test infrastructure): runs the reference's own L-BFGS
// (lib/lbfgs/lbfgs_c.cpp, lsearch.cpp, la.cpp, compiled from the reference
// tree without OpenMP, so its FP32 dot products are sequential and
// reproducible) on the reference's test functions (lib/lbfgs/functions.cpp),
// in the loop of lib/lbfgs/test1c.cpp, so tests/golden/make_golden.py can
// record the iterates.  Nothing in the product links this.
#include "functions.h"
#include "lbfgs.h"

extern "C" int ref_lbfgs_run(int func, int maxcall, float *xout, float *ftrace, int *iflag_out)
{
    int n;
    if (func == 0 || func == 1) n = 2;
    else if (func == 4) n = 50;
    else if (func == 5) n = 25;
    else n = 100;
    int m = 5, icall = 0, iflag = 0;
    float f;
    float *x = new float[n]();
    if (func == 0 || func == 1) {
        x[0] = -1.f;
        x[1] = -0.5f;
    } else if (func == 2) {
        for (int i = 0; i < n; i++) x[i] = 2.;
    } else if (func == 3) {
        for (int i = 0; i < n; i += 4) {
            x[i] = 3.f;
            x[i + 1] = -1.f;
            x[i + 3] = 1.f;
        }
    }
    float *g = new float[n]();
    float *diag = new float[n]();
    float *w = new float[n * (2 * m + 1) + 2 * m]();
    int *isave = new int[nisave]();
    float *dsave = new float[ndsave]();
    while (true) {
        if (func == 0) f = quad(x, g);
        else if (func == 1) f = camel6(x, g);
        else if (func == 2) f = rosenbrock(n, x, g);
        else if (func == 3) f = powell(n, x, g);
        else if (func == 4) f = trid(n, x, g);
        else f = rosenbrock1(n, x, g);
        ftrace[icall] = f;
        lbfgs(n, m, x, f, g, false, diag, w, iflag, isave, dsave);
        icall++;
        if (iflag <= 0 || icall >= maxcall) break;
    }
    for (int i = 0; i < n; i++) xout[i] = x[i];
    *iflag_out = iflag;
    delete[] x;
    delete[] g;
    delete[] diag;
    delete[] w;
    delete[] isave;
    delete[] dsave;
    return icall;
}

/*
 * Shared constants of the oracle (test infrastructure only).
 *
 * The finite-difference weights are the reference's truncated decimals of
 * -205/72, 8/5, -1/5, 8/315, -1/560: float literals in 2D
 * (lib/vti-fwi-2d/kernels.h:4-8) and double literals in 3D
 * (lib/vti-fwi-3d/gpu/wave3d.h:8-12).
 */
#ifndef ORC_COMMON_H
#define ORC_COMMON_H

#define C0D (-2.84722222222)
#define C1D (1.6)
#define C2D (-0.2)
#define C3D (0.02539682539)
#define C4D (-0.00178571428)

#define C0F (-2.84722222222f)
#define C1F (1.6f)
#define C2F (-0.2f)
#define C3F (0.02539682539f)
#define C4F (-0.00178571428f)

/* wave3d.h:4-6,14 */
#define PI3D 3.14159265359
#define ONETHIRD_D 0.3333333333
#define TWOTHIRD_D 0.6666666666
#define ONETHIRD_F ((float)ONETHIRD_D)
#define TWOTHIRD_F ((float)TWOTHIRD_D)
#define DAMPER3D 0.95

/* init.h:11 -- the 2D damper is a float constant */
#define DAMPER2D 0.95f

/* ---- mutation hooks ------------------------------------------------------
 * tests/test_oracle_mutants.py builds this library with -DORC_MUTANT=k, each
 * k planting one plausible mistake, and checks that the oracle's pins catch
 * every one of them.  ORC_MUTANT 0 (the default build) is the reference's
 * arithmetic; every macro below then spells the reference's expression. */
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif
/* recordData weights, injectRecord3d.cu:22-28 */
#define REC_WX (ORC_MUTANT == 1 ? ONETHIRD_D : TWOTHIRD_D)
#define REC_WZ (ORC_MUTANT == 1 ? TWOTHIRD_D : ONETHIRD_D)
/* injectResidual weights, injectRecord3d.cu:38-46 */
#define INJ_WX (ORC_MUTANT == 2 ? ONETHIRD_D : TWOTHIRD_D)
#define INJ_WZ (ORC_MUTANT == 2 ? TWOTHIRD_D : ONETHIRD_D)
/* the adjoint's first level, objFuncGradient3d_f.cu:916-920, rtm3d_f.cu:500-504 */
#define LAM0_WX (ORC_MUTANT == 3 ? ONETHIRD_D : TWOTHIRD_D)
#define LAM0_WZ (ORC_MUTANT == 3 ? TWOTHIRD_D : ONETHIRD_D)
/* gradientKernel imaging, kernel3d.cu:290-293: gc13 += lx*tz + lz*ty, gc33 += lz*tz */
#define GC13_TERM(lx, lz, ty, tz) (ORC_MUTANT == 4 ? ((lx) * (ty) + (lz) * (tz)) : ((lx) * (tz) + (lz) * (ty)))
#define GC33_TERM(lz, ty, tz) (ORC_MUTANT == 12 ? (lz) * (ty) : (lz) * (tz))
/* residual d - d_obs, objFuncGradient3d_f.cu:767-770 */
#define RESID(d, o) (ORC_MUTANT == 5 ? (o) - (d) : (d) - (o))
/* imagingKernel's receiver-side weights, kernel3d.cu:464 */
#define IMG_LWX (ORC_MUTANT == 6 ? ONETHIRD_D : TWOTHIRD_D)
#define IMG_LWZ (ORC_MUTANT == 6 ? TWOTHIRD_D : ONETHIRD_D)
/* tpower exponent, agc.cpp:103-115 */
#define TPOW_EXP(p) (ORC_MUTANT == 7 ? (p) + 1.0f : (p))
/* fine time index of the interpolated residual, objFuncGradient3d_f.cu:953 (it+1) */
#define RES_TI(it) (ORC_MUTANT == 8 ? (it) : (it) + 1)
/* imagingCrossCor, kernels.cu:414-423: 0.5 (p + q) times the slice */
#define IMG2_AQ(aq) (ORC_MUTANT == 9 ? 0 * (aq) : (aq))
/* slice interpolation weights, (1 - f) s0 + f s1 */
#define SLICE_F(f) (ORC_MUTANT == 10 ? 1 - (f) : (f))
/* extendedImagingCrossCor's offset, kernels.cu:432-433: receiver side at i - h */
#define XIMG_SH(h) (ORC_MUTANT == 11 ? -(h) : (h))

#endif

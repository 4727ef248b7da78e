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

#endif

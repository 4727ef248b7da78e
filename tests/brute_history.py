"""Stored-history restatements of the reference's 3D gradient and RTM drivers
and its 2D RTM / ODCIG drivers, in numpy FP64 (test infrastructure).

The reference never keeps the source wavefield.  In 3D it rebuilds it during
back-propagation by stepping a random-boundary run backwards
(objFuncGradient3d_f.cu:911-1078, rtm3d_f.cu:514-640); in 2D it keeps slices
every `ratio` steps and interpolates them (rtm_f.cu:95-243).  These
restatements instead run the source side forward once, keep EVERY level in
memory, and apply the imaging conditions to the stored levels.  The pairing
of levels, the weights of recording and injection, the sign of the residual,
the interpolation of the data and the damping order are read off the
reference's kernels and drivers (cited per function) and written here with
numpy array operations; nothing is shared with the oracle's C code.

Agreement with the oracle's FP64 mode therefore pins, independently of the
oracle's own body: the imaging terms of gradientKernel (kernel3d.cu:290-293)
and imagingKernel (kernel3d.cu:456-468), the 2/3-1/3 receiver weights of
recordData / injectResidual (injectRecord3d.cu:22-46), the residual's sign,
the t^p gain (agc.cpp:103-115), the time alignment of source and receiver
sides, and, in 2D, imagingCrossCor / extendedImagingCrossCor
(kernels.cu:414-436) with the slice interpolation of rtm_f / odcig_f.
"""
import numpy as np

C = (-2.84722222222, 1.6, -0.2, 0.02539682539, -0.00178571428)
TWOTHIRD = 0.6666666666  # wave3d.h:5-6
ONETHIRD = 0.3333333333
f32 = np.float32


def _d2(u, axis):
    """9-point second difference along `axis` (values at the outer 4 cells of
    that axis are meaningless and masked by the callers)."""
    out = C[0] * u
    for m in range(1, 5):
        out = out + C[m] * (np.roll(u, m, axis) + np.roll(u, -m, axis))
    return out


# ---------------------------------------------------------------------------
# 3D
# ---------------------------------------------------------------------------
class Geo3:
    """Grid scalars as the reference forms them in float (objFuncGradient3d_f.cu:374-378)."""

    def __init__(self, nz, ny, nx, d, dt):
        self.nz, self.ny, self.nx = nz, ny, nx
        self.nb = nz // 4
        self.zmax = 4 * self.nb
        dx, dy, dz = (f32(v) for v in d)
        self.dt2 = f32(dt) * f32(dt)
        self.dx2, self.dy2, self.dz2 = dx * dx, dy * dy, dz * dz
        self.sx, self.sy, self.sz = self.dt2 / self.dx2, self.dt2 / self.dy2, self.dt2 / self.dz2
        m = np.zeros((ny, nx), bool)
        m[4:ny - 4, 4:nx - 4] = True
        self.inner = m  # points forwardKernel updates in every plane


def _zwin(cz, z0, g):
    """sigma_z planes z0-4 .. z0+7 with zero ghosts outside [0, zmax)
    (forwardKernelTopBlock / BottomBlock, kernel3d.cu:76-220)."""
    w = np.zeros((12,) + cz.shape[1:])
    for k in range(12):
        z = z0 - 4 + k
        if 0 <= z < g.zmax:
            w[k] = cz[z]
    return w


def _terms3(cx, cz, z0, g, grad_form):
    """(Dxx+Dyy) sigma_x and Dzz sigma_z on the four planes z0..z0+3:
    scaled by dt2/d2 (forwardKernel, kernel3d.cu:54-61) or divided by d2
    (gradientKernel, kernel3d.cu:274-281)."""
    u = cx[z0:z0 + 4]
    lx, ly = _d2(u, 2), _d2(u, 1)
    w = _zwin(cz, z0, g)
    lz = C[0] * w[4:8] + sum(C[m] * (w[4 + m:8 + m] + w[4 - m:8 - m]) for m in range(1, 5))
    if grad_form:
        return ly / float(g.dy2) + lx / float(g.dx2), lz / float(g.dz2)
    return ly * float(g.sy) + lx * float(g.sx), lz * float(g.sz)


def _step_block3(g, ib, nxt_x, cx, nxt_z, cz, c):
    """forwardKernel on z-block ib: next = C.D(cur) + 2 cur - prev, written over
    prev at the interior x-y points of the block's 4 planes (kernel3d.cu:3-74)."""
    z0 = 4 * ib
    ty, tz = _terms3(cx, cz, z0, g, False)
    sl = slice(z0, z0 + 4)
    vx = c[0][sl] * ty + c[1][sl] * tz + 2 * cx[sl] - nxt_x[sl]
    vz = c[1][sl] * ty + c[2][sl] * tz + 2 * cz[sl] - nxt_z[sl]
    nxt_x[sl] = np.where(g.inner, vx, nxt_x[sl])
    nxt_z[sl] = np.where(g.inner, vz, nxt_z[sl])


def _damp_block3(g, ib, damping, *fields):
    """abcXYZ / abcXY (abc3d.cu:82-125): both time levels of the block times
    the float product damping[ix+iy*nx]*damping[nxy+iz], interior x-y only."""
    nxy = g.nx * g.ny
    dxy = damping[:nxy].reshape(g.ny, g.nx).astype(f32)
    for z in range(4 * ib, 4 * ib + 4):
        d = (dxy * f32(damping[nxy + z])).astype(np.float64)
        for f in fields:
            f[z] = np.where(g.inner, f[z] * d, f[z])


def _index3(loc, o, d):
    """(int)((loc - o)/d) in float, as the reference computes grid indices
    (objFuncGradient3d_f.cu:560-572)."""
    return int(f32(f32(loc) - f32(o)) / f32(d))


def geometry3(souloc, is_, recloc, stride, nx, ny, o, d):
    """Source index and block, receiver indices (true and as recorded /
    injected, i.e. block-local offset of recBlock), objFuncGradient3d_f.cu:545-576."""
    nxy = nx * ny
    nr, r0 = int(souloc[is_][3]), int(souloc[is_][4])
    rec_block = int(f32(f32(recloc.flat[2]) - f32(o[2])) / f32(d[2]) / f32(4))
    rec_true, rec_eff = [], []
    for ir in range(r0, r0 + nr):
        r = recloc.reshape(-1, stride)[ir]
        ix, iy, iz = (_index3(r[k], o[k], d[k]) for k in range(3))
        rec_true.append(ix + iy * nx + iz * nxy)
        rec_eff.append(ix + iy * nx + (iz % 4) * nxy + 4 * rec_block * nxy)
    sx, sy, sz = (_index3(souloc[is_][k], o[k], d[k]) for k in range(3))
    return dict(nr=nr, r0=r0, rec_block=rec_block, rec_true=np.array(rec_true), rec_eff=np.array(rec_eff),
                sou=sx + sy * nx + sz * nxy, sou_block=sz // 4)


def forward3(g, c, w, geo, nlev, damping=None, sts=None, nnt=None, keep=False):
    """Levels 1..nlev of one source (objFuncGradient3d_f.cu:594-735): level 1 =
    dt2*w[0] at the source, then per level and z-block: the step, the source
    dt2*w[it-1] added to the new level, the block's damping (absorbing runs);
    traces 2/3 sigma_x + 1/3 sigma_z of the damped new level at it % sts == 0.
    Returns (traces (nr, nnt) or None, list of (sigma_x, sigma_z) per level when keep)."""
    shape = (g.nz, g.ny, g.nx)
    cx, cz, px, pz = (np.zeros(shape) for _ in range(4))
    flat = lambda a: a.reshape(-1)  # noqa: E731
    flat(cx)[geo["sou"]] = float(g.dt2 * f32(w[0]))
    flat(cz)[geo["sou"]] = float(g.dt2 * f32(w[0]))
    hist = [(np.zeros(shape), np.zeros(shape)), (cx.copy(), cz.copy())] if keep else None
    tr = np.zeros((geo["nr"], nnt)) if sts else None
    for it in range(2, nlev + 1):
        for ib in range(g.nb):
            _step_block3(g, ib, px, cx, pz, cz, c)
            if ib == geo["sou_block"]:
                s = float(g.dt2 * f32(w[it - 1]))
                flat(px)[geo["sou"]] += s
                flat(pz)[geo["sou"]] += s
            if damping is not None:
                _damp_block3(g, ib, damping, px, cx, pz, cz)
        if sts and it % sts == 0 and it // sts < nnt:
            tr[:, it // sts] = TWOTHIRD * flat(px)[geo["rec_eff"]] + ONETHIRD * flat(pz)[geo["rec_eff"]]
        cx, px = px, cx
        cz, pz = pz, cz
        if keep:
            hist.append((cx.copy(), cz.copy()))
    return tr, hist


def _interp_fine(res, ti, sts):
    """interpolateResidual (injectRecord3d.cu:60-78): sample ti of the fine
    time axis from coarse traces (nr, nnt)."""
    nnt = res.shape[1]
    f = ti / sts
    i = int(f)
    if i >= nnt - 1:
        return res[:, nnt - 1]
    f -= i
    return (1.0 - f) * res[:, i] + f * res[:, i + 1]


def _adjoint3(g, c, damping, geo, src, nt, lmax, sts, on_level):
    """Receiver side of the 3D backward passes: lambda starts at level 1 with
    2/3, 1/3 of dt2*src[:, nnt-1] at the true receiver points
    (objFuncGradient3d_f.cu:916-920) and is stepped by forwardKernel with
    2/3, 1/3 of dt2 times the residual interpolated at fine index nt-s injected
    into the recBlock after its step, then damped per block
    (objFuncGradient3d_f.cu:988-1028).  on_level(s, ib, lam_cur) is called for
    s = 2..lmax before block ib steps (lam_cur as stored) and after its damping
    (lam_cur damped); returns nothing."""
    shape = (g.nz, g.ny, g.nx)
    lcx, lcz, lpx, lpz = (np.zeros(shape) for _ in range(4))
    flat = lambda a: a.reshape(-1)  # noqa: E731
    tt = float(g.dt2) * src[:, -1]
    flat(lcx)[geo["rec_true"]] = TWOTHIRD * tt
    flat(lcz)[geo["rec_true"]] = ONETHIRD * tt
    for s in range(2, lmax + 1):
        fine = _interp_fine(src, nt - s, sts)
        on_level(s, None, (lcx, lcz))
        for ib in range(g.nb):
            _step_block3(g, ib, lpx, lcx, lpz, lcz, c)
            if ib == geo["rec_block"]:
                v = float(g.dt2) * fine
                np.add.at(flat(lpx), geo["rec_eff"], TWOTHIRD * v)
                np.add.at(flat(lpz), geo["rec_eff"], ONETHIRD * v)
            _damp_block3(g, ib, damping, lpx, lcx, lpz, lcz)
            on_level(s, ib, (lcx, lcz))
        lcx, lpx = lpx, lcx
        lcz, lpz = lpz, lcz


def lmax3(nt, ngpu=1):
    """Levels the reference pipeline advances: nround*NGPU*64 + 1 (objFuncGradient3d_f.cu:539)."""
    return ((nt - 2) // (ngpu * 64)) * ngpu * 64 + 1


def gradient3(cij, padb_cij, randb_cij, w, souloc, shots, recloc, o, d, dt, nt, rate, obs, damping, tpow=0.0, ot=0.0):
    """objFuncGradientCij3d_f (objFuncGradient3d_f.cu:361-1231) with the random-
    boundary source field kept level by level.  cij: (3, nz, ny, nx) with the
    absorbing shell (pass 1); randb_cij: the same with the random shell
    (passes 2-3).  Returns (0.5*sum r^2, gradient (3, nz, ny, nx))."""
    _, nz, ny, nx = cij.shape
    g = Geo3(nz, ny, nx, d, dt)
    sts = int(round(rate / dt))
    nnt = (nt - 1) // sts + 1
    lmax = lmax3(nt)
    assert lmax == nt - 1, "stored history pairs levels exactly only when (nt-2) % 64 == 0"
    cabs = [a.astype(np.float64) for a in cij]
    crnd = [a.astype(np.float64) for a in randb_cij]
    J = 0.0
    gc = np.zeros((3, nz, ny, nx))
    for is_ in shots:
        geo = geometry3(souloc, is_, recloc, 4, nx, ny, o, d)
        tr, _ = forward3(g, cabs, w, geo, lmax, damping=damping, sts=sts, nnt=nnt)
        live = recloc.reshape(-1, 4)[geo["r0"]:geo["r0"] + geo["nr"], 3] != 0
        tr[~live] = 0.0
        if tpow != 0.0:
            t = f32(ot) + np.arange(nnt, dtype=f32) * f32(rate)
            tr = tr * np.power(t.astype(np.float64), float(f32(tpow)))
        ob = np.asarray(obs, np.float64)[nnt * geo["r0"]:nnt * (geo["r0"] + geo["nr"])].reshape(geo["nr"], nnt)
        res = tr - ob
        J += float(np.sum(res * res))
        _, hist = forward3(g, crnd, w, geo, lmax, keep=True)

        def image(s, ib, lam):
            if ib is not None:
                return
            lx, lz = lam
            sx, sz = hist[lmax + 1 - s]
            for b in range(g.nb):
                ty, tz = _terms3(sx, sz, 4 * b, g, True)
                sl = slice(4 * b, 4 * b + 4)
                gc[0][sl] += np.where(g.inner, lx[sl] * ty, 0.0)
                gc[2][sl] += np.where(g.inner, lz[sl] * tz, 0.0)
                gc[1][sl] += np.where(g.inner, lx[sl] * tz + lz[sl] * ty, 0.0)

        _adjoint3(g, crnd, damping, geo, res, nt, lmax, sts, image)
    return J / 2, gc


def rtm3(c, w, souloc, shots, recloc3, o, d, dt, nt, rate, obs, damping):
    """rtmCij3d_f (rtm3d_f.cu:161-830) with the source field kept level by
    level: image += (2/3 sx + 1/3 sz)(2/3 lx + 1/3 lz) per block after the
    receiver side's damping (rtm3d_f.cu:585-610).  recloc stride 3."""
    nz, ny, nx = c[0].shape
    g = Geo3(nz, ny, nx, d, dt)
    sts = int(round(rate / dt))
    nnt = (nt - 1) // sts + 1
    lmax = lmax3(nt)
    assert lmax == nt - 1
    cc = [a.astype(np.float64) for a in c]
    img = np.zeros((nz, ny, nx))
    for is_ in shots:
        geo = geometry3(souloc, is_, recloc3, 3, nx, ny, o, d)
        _, hist = forward3(g, cc, w, geo, lmax, keep=True)
        ob = np.asarray(obs, np.float64)[nnt * geo["r0"]:nnt * (geo["r0"] + geo["nr"])].reshape(geo["nr"], nnt)

        def image(s, ib, lam):
            if ib is None:
                return
            lx, lz = lam
            sx, sz = hist[lmax + 1 - s]
            sl = slice(4 * ib, 4 * ib + 4)
            v = (TWOTHIRD * sx[sl] + ONETHIRD * sz[sl]) * (TWOTHIRD * lx[sl] + ONETHIRD * lz[sl])
            img[sl] += np.where(g.inner, v, 0.0)

        _adjoint3(g, cc, damping, geo, ob, nt, lmax, sts, image)
    return img


# ---------------------------------------------------------------------------
# 2D (rtm_f / odcig_f)
# ---------------------------------------------------------------------------
def _step2(p0, q0, p1, q1, c, dx2, dz2, dt2):
    """forwardCD (kernels.cu:452-498): p0 <- 2 p1 + dt2 (c11 Dxx p1 + c13 Dzz q1) - p0,
    q0 likewise with c13, c33, at [4, nnx-4) x [4, nnz-4)."""
    tp = _d2(p1, 1) / dx2
    tq = _d2(q1, 0) / dz2
    vp = 2 * p1 + dt2 * (c[0] * tp + c[1] * tq) - p0
    vq = 2 * q1 + dt2 * (c[1] * tp + c[2] * tq) - q0
    m = np.zeros(p0.shape, bool)
    m[4:-4, 4:-4] = True
    p0[m] = vp[m]
    q0[m] = vq[m]


def _dipole(p, q, i, nnx, s):
    """injectDipoleSource / injectDipoleData (kernels.cu:232-257): +s at i, -s one row below."""
    fp, fq = p.reshape(-1), q.reshape(-1)
    np.add.at(fp, i, s)
    np.add.at(fq, i, s)
    np.add.at(fp, i + nnx, -s)
    np.add.at(fq, i + nnx, -s)


def rtm2(data, c, w, sloc, rloc, taper, nnx, nnz, nt, dx, dz, dt, ratio, ntneg, nhx):
    """rtm_f (nhx < 0, rtm_f.cu:13-269) or odcig_f (odcig_f.cu:12-247) with every
    level of the source side kept.  The source slice used at receiver level it
    is (1-f) S[k] + f S[k+1], k = floor(it/ratio), f = it/ratio - k, where
    S[k] = (p+q)/2 of source level k*ratio right after its taper (S[0] = 0; for
    ratio 1 also level 1); the data injected at level it is
    (1-f) row[k] + f row[k+1], k = floor((it-ntneg+1)/ratio), rows past the last
    one zero, as the reference's buffer rotations give them."""
    f32_ = np.float32
    dx2, dz2, dt2 = float(f32_(dx) * f32_(dx)), float(f32_(dz) * f32_(dz)), float(f32_(dt) * f32_(dt))
    cc = [np.asarray(a, np.float64).reshape(nnz, nnx) for a in c]
    tap = np.asarray(taper, np.float64).reshape(nnz, nnx)
    inner = np.zeros((nnz, nnx), bool)
    inner[4:-4, 4:-4] = True
    tp_in = np.where(inner, tap, 1.0)  # abc multiplies the updated interior only (kernels.cu:293-315)
    data = np.asarray(data, np.float64).reshape(-1, len(rloc))
    nnt_data = data.shape[0]
    nnt = (nt - 1) // ratio + 1
    odcig = nhx >= 0
    img = np.zeros((2 * nhx + 1, nnz, nnx) if odcig else (nnz, nnx))
    for s_ in sloc:
        isrc = int(s_[0]) + int(s_[1]) * nnx
        nr, r0 = int(s_[2]), int(s_[3])
        ridx = rloc[r0:r0 + nr, 0] + rloc[r0:r0 + nr, 1] * nnx
        # source side, every level
        p0, q0, p1, q1 = (np.zeros((nnz, nnx)) for _ in range(4))
        _dipole(p1, q1, isrc, nnx, float(f32_(dt2) * f32_(w[0])))
        p1 *= tp_in
        q1 *= tp_in
        S = np.zeros((nnt + 1, nnz, nnx))
        if ratio == 1:
            S[1] = np.where(inner, 0.5 * (p1 + q1), 0.0)
        for it in range(2, nt):
            _step2(p0, q0, p1, q1, cc, dx2, dz2, dt2)
            amp = float(f32_(w[it - 1])) if odcig else float(f32_(dt2) * f32_(w[it - 1]))
            _dipole(p0, q0, isrc, nnx, amp)
            for a in (p1, q1, p0, q0):
                a *= tp_in
            if it % ratio == 0:
                S[it // ratio] = np.where(inner, 0.5 * (p0 + q0), 0.0)
            p0, p1 = p1, p0
            q0, q1 = q1, q0

        def slice_at(it):
            k = it // ratio
            f = it / ratio - k
            return (1 - f) * S[k] + f * S[k + 1]

        def rows_at(it):
            m = it - ntneg + 1
            k = m // ratio
            f = m / ratio - k
            a = data[k, r0:r0 + nr] if k < nnt_data else 0.0
            b = data[k + 1, r0:r0 + nr] if k + 1 < nnt_data else 0.0
            return 0.5 * ((1 - f) * a + f * b)

        def image(ap, aq, it):
            d = slice_at(it)
            if not odcig:
                img[inner] += (0.5 * (ap + aq) * d)[inner]
                return
            a = 0.5 * (ap + aq)
            for h in range(-nhx, nhx + 1):
                # image[h][i] += a(i-h) d(i+h) at ix in [nhx, nnx-nhx)
                lo, hi = nhx, nnx - nhx
                img[h + nhx][:, lo:hi] += a[:, lo - h:hi - h] * d[:, lo + h:hi + h]

        # receiver side, backwards
        p0, q0, p1, q1 = (np.zeros((nnz, nnx)) for _ in range(4))
        _dipole(p0, q0, ridx, nnx, dt2 * 0.5 * data[nnt_data - 1, r0:r0 + nr])
        p0 *= tp_in
        q0 *= tp_in
        image(p0, q0, nt - 2)
        for it in range(nt - 3, -1, -1):
            _step2(p1, q1, p0, q0, cc, dx2, dz2, dt2)
            if it >= ntneg:
                _dipole(p1, q1, ridx, nnx, dt2 * rows_at(it))
            for a in (p1, q1, p0, q0):
                a *= tp_in
            image(p1, q1, it)
            p0, p1 = p1, p0
            q0, q1 = q1, q0
    return img

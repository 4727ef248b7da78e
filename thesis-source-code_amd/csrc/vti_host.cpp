// Host-side helpers: see vti_host.h.  Compiled with -ffp-contract=off so
// every expression rounds exactly as the reference's host code does.
#include "vti_host.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <iterator>
#include <set>
#include <vector>

namespace vti {

namespace {
constexpr double kPi3D = 3.14159265359;  // wave3d.h:4
constexpr double kDamper3D = 0.95;       // wave3d.h:14
constexpr float kDamper2D = 0.95f;       // init.h:11
}  // namespace

void init_damping3d(float *damping, int nx, int ny, int nz, int npad)
{
    auto profile = [npad](int i, int n) {
        float dist;
        if (i < npad) dist = (float)(npad - i);
        else if (i > n - npad) dist = (float)(i - n + npad);
        else dist = 0.f;
        return (float)(kDamper3D + (1. - kDamper3D) * std::cos(kPi3D * (double)dist / npad));
    };
    for (int iy = 0; iy < ny; iy++) {
        const float dy = profile(iy, ny);
        for (int ix = 0; ix < nx; ix++) damping[ix + (size_t)iy * nx] = profile(ix, nx) * dy;
    }
    const size_t nxy = (size_t)nx * ny;
    for (int iz = 0; iz < nz; iz++) damping[nxy + iz] = profile(iz, nz);
}

void taper_profile2d(float *t0, int npad)
{
    // init.h:9 defines pi as 4.f*atan(1.f): the float value of pi, widened.
    const double pi = (double)(4.f * std::atan(1.f));
    for (int i = 0; i < npad; ++i)
        t0[i] = (float)((double)kDamper2D + (1. - (double)kDamper2D) * std::cos(pi * (double)(float)(npad - 1 - i) / npad));
}

void init_taper2d(float *taper, int nx, int nz, int npad)
{
    std::vector<float> t0(npad > 0 ? npad : 1);
    taper_profile2d(t0.data(), npad);
    const int nnx = nx + 2 * npad, nnz = nz + 2 * npad;
    for (size_t i = 0; i < (size_t)nnx * nnz; ++i) taper[i] = 1.f;
    for (int i = 0; i < nnz; ++i) {
        float *row = taper + (size_t)i * nnx;
        for (int j = 0; j < npad; ++j) row[j] = row[j] * t0[j];
        for (int j = 0; j < npad; ++j) row[npad + nx + j] = row[npad + nx + j] * t0[npad - 1 - j];
    }
    for (int i = 0; i < npad; ++i) {
        float *a = taper + (size_t)i * nnx;
        float *b = taper + (size_t)(i + npad + nz) * nnx;
        for (int j = 0; j < nnx; ++j) a[j] = a[j] * t0[i];
        for (int j = 0; j < nnx; ++j) b[j] = b[j] * t0[npad - 1 - i];
    }
}

void vepsdel_to_cij(float *c11, float *c13, float *c33, const float *v, const float *eps, const float *del, float v0,
                    float eps0, float del0, size_t n)
{
    for (size_t i = 0; i < n; ++i) {
        c33[i] = v0 * v0 * v[i] * v[i];
        c11[i] = (float)((double)c33[i] * (1. + (double)(2 * eps0 * eps[i])));
        c13[i] = (float)((double)c33[i] * std::sqrt(1. + (double)(2 * del0 * del[i])));
    }
}

void dvepsdel_to_dcij(float *c11, float *c13, float *c33, float *dc11, float *dc13, float *dc33, const float *v,
                      const float *eps, const float *del, const float *dv, const float *deps, const float *ddel,
                      float v0, float eps0, float del0, size_t n)
{
    for (size_t i = 0; i < n; ++i) {
        float vv = v0 * v[i], ee = eps0 * eps[i], dd = del0 * del[i];
        float vv2 = vv * vv;
        ee = (float)(1. + (double)(2 * ee));
        dd = (float)std::sqrt(1. + (double)(2 * dd));
        c11[i] = vv2 * ee;
        c13[i] = vv2 * dd;
        c33[i] = vv2;
        dc33[i] = 2 * vv * dv[i] * v0;
        dc11[i] = dc33[i] * ee + 2 * vv2 * deps[i] * eps0;
        dc13[i] = dc33[i] * dd + vv2 / dd * ddel[i] * del0;
    }
}

void gradcij_to_gradvepsdel(float *gv, float *geps, float *gdel, const float *gc11, const float *gc13,
                            const float *gc33, const float *v, const float *eps, const float *del, float v0, float eps0,
                            float del0, size_t n)
{
    for (size_t i = 0; i < n; ++i) {
        float tv = v0 * v[i];
        float teps = (float)(1. + (double)(2 * eps0 * eps[i]));
        float tdel = (float)std::sqrt(1. + (double)(2 * del0 * del[i]));
        gv[i] = 2 * tv * (gc11[i] * teps + gc13[i] * tdel + gc33[i]) * v0;
        tv = tv * tv;
        geps[i] = 2 * tv * gc11[i] * eps0;
        gdel[i] = tv / tdel * gc13[i] * del0;
    }
}

long long boundary_size(int nx, int ny, int nz, int npad)
{
    return (long long)nx * ny * nz - (long long)(nx - 2 * npad) * (ny - 2 * npad) * (nz - 2 * npad);
}

namespace {
// Walks the shell in the packing order of boundary.cpp:10-46: the top slab,
// then for each interior plane the npad front rows, [npad | npad] of every
// interior row, the npad back rows, and finally the bottom slab.
template <class F>
void walk_shell(int nx, int ny, int nz, int npad, F &&seg)
{
    long long cv = 0;
    long long n = (long long)nx * ny * npad;
    seg(cv, n);
    cv += n;
    for (int iz = 0; iz < nz - 2 * npad; iz++) {
        for (int iy = 0; iy < npad; iy++) {
            seg(cv, nx);
            cv += nx;
        }
        for (int iy = 0; iy < ny - 2 * npad; iy++) {
            seg(cv, npad);
            cv += nx - npad;
            seg(cv, npad);
            cv += npad;
        }
        for (int iy = 0; iy < npad; iy++) {
            seg(cv, nx);
            cv += nx;
        }
    }
    seg(cv, n);
}
}  // namespace

void get_boundary(float *packed, const float *v, int nx, int ny, int nz, int npad)
{
    long long cb = 0;
    walk_shell(nx, ny, nz, npad, [&](long long cv, long long n) {
        std::memcpy(packed + cb, v + cv, sizeof(float) * n);
        cb += n;
    });
}

void put_boundary(const float *packed, float *v, int nx, int ny, int nz, int npad)
{
    long long cb = 0;
    walk_shell(nx, ny, nz, npad, [&](long long cv, long long n) {
        std::memcpy(v + cv, packed + cb, sizeof(float) * n);
        cb += n;
    });
}

void zero_boundary(float *v, int nx, int ny, int nz, int npad)
{
    walk_shell(nx, ny, nz, npad, [&](long long cv, long long n) { std::memset(v + cv, 0, sizeof(float) * n); });
}

void ricker(float *wavelet, float freq, int nt, float dt, float ot, float tdelay, float scalefactor)
{
    const double pi = 4. * std::atan(1.);
    const float pi2f2 = (float)(pi * pi * freq * freq);
    for (int it = 0; it < nt; ++it) {
        float t = it * dt + ot - tdelay;
        float pi2f2t2 = pi2f2 * t * t;
        // ricker.cpp calls exp on a float argument, which resolves to the float overload
        wavelet[it] = (float)((1. - 2. * (double)pi2f2t2) * (double)std::exp(-pi2f2t2));
        wavelet[it] *= scalefactor;
    }
}

// ---- glibc random(): srandom_r / random_r with the default TYPE_3 state ----
void GlibcRand::seed_(unsigned seed)
{
    if (seed == 0) seed = 1;
    s_[0] = (int32_t)seed;
    int32_t word = (int32_t)seed;
    for (int i = 1; i < 31; ++i) {
        // state[i] = 16807 * state[i-1] % 2147483647 without overflowing 31 bits
        const long hi = word / 127773, lo = word % 127773;
        word = (int32_t)(16807 * lo - 2836 * hi);
        if (word < 0) word += 2147483647;
        s_[i] = word;
    }
    f_ = 3;
    r_ = 0;
    for (int i = 0; i < 310; ++i) next();
}

int GlibcRand::next()
{
    const uint32_t val = (uint32_t)s_[f_] + (uint32_t)s_[r_];
    s_[f_] = (int32_t)val;
    const int result = (int)(val >> 1);
    if (++f_ >= 31) {
        f_ = 0;
        ++r_;
    } else if (++r_ >= 31) {
        r_ = 0;
    }
    return result;
}

// ---- RandomMY (randomBound.h:51-392) --------------------------------------
namespace {

constexpr int kRandMax = 2147483647;  // RAND_MAX of glibc
constexpr int kRndMax = 50000;        // size of the pre-drawn uniform block

struct RndBlock {
    GlibcRand &g;
    // one extra slot: the reference's pre-increment read rnd[++irand] can
    // touch rnd[50000], one past the block, which is zero in the reference's
    // process (the block is a fresh zero-filled mapping; see DESIGN notes).
    std::vector<float> v = std::vector<float>(kRndMax + 1, 0.f);
    int i = 1;
    explicit RndBlock(GlibcRand &gg) : g(gg) {}
    void reset()
    {
        for (int k = 0; k < kRndMax; ++k) v[k] = (float)((1. * g.next()) / kRandMax);
        i = 0;
    }
};

}  // namespace

long long random_boundary(float *field, float bval, float eval, int n1, int n2, int n3, int t1, int t2, int t3,
                          float pctG, int seed)
{
    if (n1 <= 0 || n2 <= 0 || n3 <= 0) return -1;
    GlibcRand rng((unsigned)seed);
    const int ndir = n3 == 1 ? 8 : 26;
    int dir[26][3];
    float pdf[26];
    {
        int cnt = 0;
        if (ndir == 8) {
            const float s1 = (float)std::exp(-1.), s2 = (float)std::exp(-2.);
            for (int a = -1; a <= 1; ++a)
                for (int b = -1; b <= 1; ++b) {
                    if (a == 0 && b == 0) continue;
                    dir[cnt][0] = a;
                    dir[cnt][1] = b;
                    dir[cnt][2] = 0;
                    pdf[cnt] = (a != 0 && b != 0) ? s2 : s1;
                    ++cnt;
                }
        } else {
            for (int a = -1; a <= 1; ++a)
                for (int b = -1; b <= 1; ++b)
                    for (int c = -1; c <= 1; ++c) {
                        if (a == 0 && b == 0 && c == 0) continue;
                        dir[cnt][0] = a;
                        dir[cnt][1] = b;
                        dir[cnt][2] = c;
                        pdf[cnt] = 1.f;
                        ++cnt;
                    }
        }
    }
    const long long model = (long long)n1 * n2 * n3;
    const long long bsize = model - (long long)(n1 - 2 * t1) * (n2 - 2 * t2) * std::max(n3 - 2 * t3, 1);
    if (bsize <= 0) return -1;
    std::vector<int> label(model);
    std::vector<int> b2g(bsize), g2b(model, 0);
    std::vector<float> dist(bsize + 1), start(bsize + 2, 0.f), prop(bsize + 1, 1.f);
    std::vector<char> used(bsize, 0);
    std::vector<int> minrad(bsize, 1);

    // distance of every shell point to the interior box (initialDistanceFunction)
    long long cb = 0, cg = 0;
    for (int i3 = 0; i3 < n3; i3++) {
        int in3 = std::min(std::max(i3, t3), n3 - t3 - 1);
        in3 = std::min(std::max(in3, 0), n3 - 1);
        const float d3 = (float)((i3 - in3) * (i3 - in3));
        for (int i2 = 0; i2 < n2; i2++) {
            const int in2 = std::min(std::max(i2, t2), n2 - t2 - 1);
            const float d2 = (float)((i2 - in2) * (i2 - in2));
            for (int i1 = 0; i1 < n1; i1++) {
                const int in1 = std::min(std::max(i1, t1), n1 - t1 - 1);
                const float d1 = (float)((i1 - in1) * (i1 - in1));
                const float dd = std::sqrt(d1 + d2 + d3);
                if (cb <= bsize) dist[cb] = dd;
                label[cg] = -1;
                if (dd > 0.01f) {
                    label[cg] = 0;
                    b2g[cb] = (int)cg;
                    g2b[cg] = (int)cb;
                    cb++;
                }
                cg++;
            }
        }
    }
    if (cb != bsize || cg != model) return -1;

    RndBlock rnd(rng);
    rnd.reset();
    long long left = bsize;
    int cluster_id = 0;
    int residual_dir[26];
    float cdf[26];
    bool tried[26];
    const long long n12 = (long long)n1 * n2;

    while (left > 0) {
        // firstUnusedPoint
        int iloc = (int)std::round((float)(bsize - 1) * rnd.v[rnd.i]);
        rnd.i++;
        if (rnd.i >= kRndMax) rnd.reset();
        if (used[iloc]) {
            int idist = minrad[iloc];
            for (;;) {
                const int lo = std::max(iloc - idist, 0);
                if (!used[lo]) {
                    minrad[iloc] = idist - 1;
                    iloc = lo;
                    break;
                }
                const int hi = (int)std::min((long long)iloc + idist, bsize - 1);
                if (!used[hi]) {
                    minrad[iloc] = idist - 1;
                    iloc = hi;
                    break;
                }
                idist++;
            }
        }
        std::set<int> cluster;
        cluster.insert(iloc);
        used[iloc] = 1;
        cluster_id++;
        long long gp = b2g[iloc];
        int i1 = (int)(gp % n1), i2 = (int)((gp / n1) % n2), i3 = (int)(gp / n12);
        label[gp] = cluster_id;
        start[cluster_id] = dist[iloc];
        left--;

        bool grow = true, reselect = false;
        while (grow && left > 0 && !cluster.empty()) {
            if (!reselect && rnd.v[rnd.i] > pctG) {
                grow = false;
                ++rnd.i;
                if (rnd.i >= kRndMax - 10) rnd.reset();
                continue;
            }
            reselect = false;
            const int shift = (int)std::round((float)(cluster.size() - 1) * rnd.v[++rnd.i]);
            auto it = cluster.begin();
            std::advance(it, shift);
            iloc = *it;
            gp = b2g[iloc];
            i1 = (int)(gp % n1);
            i2 = (int)((gp / n1) % n2);
            i3 = (int)(gp / n12);

            int nleft = 0;
            cdf[0] = 0.f;
            for (int k = 0; k < ndir; k++) {
                tried[k] = true;
                const int a = i1 + dir[k][0], b = i2 + dir[k][1], c = i3 + dir[k][2];
                if (a >= 0 && b >= 0 && c >= 0 && a < n1 && b < n2 && c < n3) {
                    const long long coord = (long long)c * n12 + (long long)b * n1 + a;
                    if (label[coord] == 0) {
                        tried[k] = false;
                        residual_dir[nleft] = k;
                        cdf[nleft] = cdf[std::max(nleft - 1, 0)] + pdf[k];
                        nleft++;
                    }
                }
            }
            if (nleft == 0) {
                reselect = true;
                cluster.erase(iloc);
                continue;
            }
            bool tried_all = false, search = true;
            while (!tried_all && search) {
                const float prob = cdf[nleft - 1] * rnd.v[rnd.i++];
                if (rnd.i >= kRndMax) rnd.reset();
                int idir = 0;
                while (cdf[idir] < prob) idir++;
                idir = residual_dir[idir];
                if (!tried[idir]) {
                    tried[idir] = true;
                    const int a = i1 + dir[idir][0], b = i2 + dir[idir][1], c = i3 + dir[idir][2];
                    if (a >= 0 && b >= 0 && c >= 0 && a < n1 && b < n2 && c < n3) {
                        const long long coord = (long long)c * n12 + (long long)b * n1 + a;
                        if (label[coord] == 0) {
                            label[coord] = cluster_id;
                            used[g2b[coord]] = 1;
                            left--;
                            search = false;
                            i1 = a;
                            i2 = b;
                            i3 = c;
                            cluster.insert(g2b[coord]);
                        }
                    }
                }
                tried_all = true;
                for (int k = 0; k < ndir; k++)
                    if (!tried[k]) {
                        tried_all = false;
                        break;
                    }
                if (tried_all && search) grow = false;
            }
        }
    }

    // per-cluster factors; index 0 holds the reference's never-written start[0]
    // (zero), and the last cluster's factor stays 1 (loop bound i < clusterId)
    const float dmax = (float)std::max(std::max(t1, t2), t3);
    for (int i = 0; i < cluster_id; i++) {
        const float pct = start[i] / dmax;
        const float dbase = (float)((1. - (double)pct) * bval + (double)(pct * eval));
        if (pct == 0) {
            prop[i] = 1;
            continue;
        }
        for (;;) {
            float r = (float)rng.next() / (float)kRandMax;
            r = (float)(((double)r - 0.5) * (double)pct + (double)dbase);
            if (eval > bval) {
                if ((double)r > 1. && r < eval * 3) {
                    prop[i] = r;
                    break;
                }
            } else {
                if ((double)r > (double)eval / 3. && (double)r < 1.) {
                    prop[i] = r;
                    break;
                }
            }
        }
    }
    for (long long loc = 0; loc < model; ++loc)
        if (label[loc] > 0) field[loc] = field[loc] * prop[label[loc]];
    return cluster_id;
}

// checkEpsDel, check.cpp:22-40: FP32 throughout, as the reference computes it
long long check_eps_del(float *eps, float *del, float eps0, float del0, size_t n, float *m)
{
    long long count = 0;
    for (size_t i = 0; i < n; ++i) {
        m[i] = 0.f;
        const float d = del0 * del[i], e = eps0 * eps[i];
        if (d == e) {
            m[i] = 1.f;
        } else if (d > e) {
            const float t = (eps0 * del[i] + del0 * eps[i]) / (eps0 * eps0 + del0 * del0);
            eps[i] = t * del0;
            del[i] = t * eps0;
            m[i] = 2.f;
            ++count;
        }
    }
    return count;
}

}  // namespace vti

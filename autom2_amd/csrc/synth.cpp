// synth.cpp -- deterministic synthetic tilt series (libtrec_synth.so).
//
// A phantom of random ellipsoids plus dense spheres ("gold beads"),
// projected with exact line integrals in TiltRec's detector convention:
// pixel (i, j) of tilt theta integrates the line
//     cos(theta) (x - X/2) - sin(theta) (z - T/2) = i - X/2,  y = j
// through the slab z in [-1/2, T - 1/2].  Then Gaussian noise with sigma =
// noise_frac * (mean projection value) and a per-tilt shift / in-plane
// rotation resampled bilinearly about the image centre (border = the image
// mean), whose true values come back in `xf` as IMOD .xf rows
// (cos phi, -sin phi, sin phi, cos phi, dx, dy).  Everything derives from
// one SplitMix64 stream, so a seed reproduces the series bit for bit.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uni() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
    double uni(double a, double b) { return a + (b - a) * uni(); }
    double gauss()
    {
        double u1 = uni(), u2 = uni();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

struct Body {
    double c[3];     // centre (x, y, z), voxel units
    double ax[3];    // semi-axes
    double R[3][3];  // columns: body axes in world coordinates
    double rho;      // density
};

void random_rotation(Rng &r, double R[3][3])
{
    // uniform unit quaternion (Shoemake)
    const double u1 = r.uni(), u2 = r.uni(), u3 = r.uni();
    const double a = std::sqrt(1 - u1), b = std::sqrt(u1);
    const double q0 = a * std::sin(2 * M_PI * u2), q1 = a * std::cos(2 * M_PI * u2);
    const double q2 = b * std::sin(2 * M_PI * u3), q3 = b * std::cos(2 * M_PI * u3);
    R[0][0] = 1 - 2 * (q2 * q2 + q3 * q3);
    R[0][1] = 2 * (q1 * q2 - q0 * q3);
    R[0][2] = 2 * (q1 * q3 + q0 * q2);
    R[1][0] = 2 * (q1 * q2 + q0 * q3);
    R[1][1] = 1 - 2 * (q1 * q1 + q3 * q3);
    R[1][2] = 2 * (q2 * q3 - q0 * q1);
    R[2][0] = 2 * (q1 * q3 - q0 * q2);
    R[2][1] = 2 * (q2 * q3 + q0 * q1);
    R[2][2] = 1 - 2 * (q1 * q1 + q2 * q2);
}

// Half-width of the body's support along unit direction n.
double support(const Body &b, const double n[3])
{
    double s = 0;
    for (int k = 0; k < 3; ++k) {
        double d = 0;
        for (int m = 0; m < 3; ++m) d += n[m] * b.R[m][k];
        s += d * d * b.ax[k] * b.ax[k];
    }
    return std::sqrt(s);
}

// Length of the part of line p0 + t d (|d| = 1) inside the body, t in [t0, t1].
double chord(const Body &b, const double p0[3], const double d[3], double t0, double t1)
{
    double q0[3], qd[3];
    for (int k = 0; k < 3; ++k) {
        double a = 0, e = 0;
        for (int m = 0; m < 3; ++m) {
            a += (p0[m] - b.c[m]) * b.R[m][k];
            e += d[m] * b.R[m][k];
        }
        q0[k] = a / b.ax[k];
        qd[k] = e / b.ax[k];
    }
    const double A = qd[0] * qd[0] + qd[1] * qd[1] + qd[2] * qd[2];
    const double B = 2 * (q0[0] * qd[0] + q0[1] * qd[1] + q0[2] * qd[2]);
    const double C = q0[0] * q0[0] + q0[1] * q0[1] + q0[2] * q0[2] - 1;
    const double disc = B * B - 4 * A * C;
    if (disc <= 0) return 0;
    const double sq = std::sqrt(disc);
    double ta = (-B - sq) / (2 * A), tb = (-B + sq) / (2 * A);
    ta = std::max(ta, t0);
    tb = std::min(tb, t1);
    return tb > ta ? tb - ta : 0;
}

}  // namespace

extern "C" {

// stack: [P][Y][X] float out; xf: [P][6] float out (may be NULL); volume:
// optional [Y][T][X] voxelised phantom (density at voxel centres), may be NULL.
// scale multiplies every length of the recipe (semi-axes 4..30, bead radius 3).
// Returns 0 on success.
int trec_synth_series(int X, int Y, int T, int P, const float *angles_deg, int n_ellipsoids, int n_beads,
                      double scale, double noise_frac, double shift_max, double rot_max_deg, uint64_t seed,
                      float *stack, float *xf, float *volume)
{
    if (X < 2 || Y < 1 || T < 1 || P < 1 || !angles_deg || !stack) return 1;
    Rng rng(seed);
    std::vector<Body> bodies;
    for (int k = 0; k < n_ellipsoids + n_beads; ++k) {
        Body b;
        const bool bead = k >= n_ellipsoids;
        b.c[0] = rng.uni(0.1, 0.9) * X;
        b.c[1] = rng.uni(0.1, 0.9) * Y;
        b.c[2] = rng.uni(0.1, 0.9) * T;
        for (int m = 0; m < 3; ++m) b.ax[m] = bead ? 3.0 * scale : rng.uni(4.0, 30.0) * scale;
        random_rotation(rng, b.R);
        b.rho = bead ? 5.0 : rng.uni(0.2, 1.0);
        bodies.push_back(b);
    }
    const double ox = X * 0.5, oz = T * 0.5;
    const size_t img = (size_t)X * Y;

#pragma omp parallel for schedule(dynamic)
    for (int p = 0; p < P; ++p) {
        float *out = stack + (size_t)p * img;
        std::memset(out, 0, sizeof(float) * img);
        const double th = angles_deg[p] * M_PI / 180.0;
        const double ct = std::cos(th), st = std::sin(th);
        const double d[3] = {st, 0.0, ct};     // ray direction
        const double n[3] = {ct, 0.0, -st};    // detector x direction
        const double ey[3] = {0.0, 1.0, 0.0};
        for (const Body &b : bodies) {
            const double uc = ox + ct * (b.c[0] - ox) - st * (b.c[2] - oz);
            const double hu = support(b, n), hy = support(b, ey);
            const int i0 = std::max(0, (int)std::floor(uc - hu)), i1 = std::min(X - 1, (int)std::ceil(uc + hu));
            const int j0 = std::max(0, (int)std::floor(b.c[1] - hy)), j1 = std::min(Y - 1, (int)std::ceil(b.c[1] + hy));
            for (int j = j0; j <= j1; ++j)
                for (int i = i0; i <= i1; ++i) {
                    const double p0[3] = {ox + (i - ox) * ct, (double)j, oz - (i - ox) * st};
                    // slab z in [-0.5, T - 0.5]
                    double t0 = -1e30, t1 = 1e30;
                    if (std::fabs(ct) > 1e-12) {
                        double ta = (-0.5 - p0[2]) / ct, tb = (T - 0.5 - p0[2]) / ct;
                        t0 = std::min(ta, tb);
                        t1 = std::max(ta, tb);
                    }
                    const double L = chord(b, p0, d, t0, t1);
                    if (L > 0) out[(size_t)j * X + i] += (float)(L * b.rho);
                }
        }
    }

    double mean = 0;
    for (size_t k = 0; k < (size_t)P * img; ++k) mean += stack[k];
    mean /= (double)P * img;
    const double sigma = noise_frac * mean;
    // noise and per-tilt parameters drawn in a fixed order
    std::vector<double> prm(6 * (size_t)P);
    for (int p = 0; p < P; ++p) {
        float *o = stack + (size_t)p * img;
        for (size_t k = 0; k < img; ++k) o[k] += (float)(sigma * rng.gauss());
        const double phi = rng.uni(-rot_max_deg, rot_max_deg) * M_PI / 180.0;
        prm[6 * p + 0] = std::cos(phi);
        prm[6 * p + 1] = -std::sin(phi);
        prm[6 * p + 2] = std::sin(phi);
        prm[6 * p + 3] = std::cos(phi);
        prm[6 * p + 4] = rng.uni(-shift_max, shift_max);
        prm[6 * p + 5] = rng.uni(-shift_max, shift_max);
        if (xf)
            for (int m = 0; m < 6; ++m) xf[6 * p + m] = (float)prm[6 * p + m];
    }
    if (shift_max > 0 || rot_max_deg > 0) {
#pragma omp parallel for schedule(dynamic)
        for (int p = 0; p < P; ++p) {
            float *o = stack + (size_t)p * img;
            std::vector<float> src(o, o + img);
            double m = 0;
            for (size_t k = 0; k < img; ++k) m += src[k];
            const float fill = (float)(m / img);
            const double *q = &prm[6 * p];
            const double cx = X * 0.5, cy = Y * 0.5;
            // aligned = A * raw + shift about the centre; the raw image is
            // sampled at A^-1 (aligned - centre - shift) + centre
            const double det = q[0] * q[3] - q[1] * q[2];
            const double i00 = q[3] / det, i01 = -q[1] / det, i10 = -q[2] / det, i11 = q[0] / det;
            for (int j = 0; j < Y; ++j)
                for (int i = 0; i < X; ++i) {
                    const double ax = i - cx - q[4], ay = j - cy - q[5];
                    const double sx = i00 * ax + i01 * ay + cx, sy = i10 * ax + i11 * ay + cy;
                    const int x0 = (int)std::floor(sx), y0 = (int)std::floor(sy);
                    const double fx = sx - x0, fy = sy - y0;
                    double v = 0;
                    for (int dy = 0; dy < 2; ++dy)
                        for (int dx = 0; dx < 2; ++dx) {
                            const int xx = x0 + dx, yy = y0 + dy;
                            const double w = (dx ? fx : 1 - fx) * (dy ? fy : 1 - fy);
                            const float s = (xx >= 0 && xx < X && yy >= 0 && yy < Y) ? src[(size_t)yy * X + xx] : fill;
                            v += w * s;
                        }
                    o[(size_t)j * X + i] = (float)v;
                }
        }
    }

    if (volume) {
#pragma omp parallel for schedule(dynamic)
        for (int y = 0; y < Y; ++y)
            for (int z = 0; z < T; ++z)
                for (int x = 0; x < X; ++x) {
                    double v = 0;
                    for (const Body &b : bodies) {
                        double r = 0;
                        for (int k = 0; k < 3; ++k) {
                            const double e = ((x - b.c[0]) * b.R[0][k] + (y - b.c[1]) * b.R[1][k] + (z - b.c[2]) * b.R[2][k]) / b.ax[k];
                            r += e * e;
                        }
                        if (r <= 1) v += b.rho;
                    }
                    volume[((size_t)y * T + z) * X + x] = (float)v;
                }
    }
    return 0;
}

}  // extern "C"

/*
 * synth.c -- seeded procedural RGB-D renderer for the odometry workloads.
 *
 * Scene: the inside of an axis-aligned room (6 planes) holding axis-aligned
 * boxes.  Every surface carries a solid (3-D) Perlin texture, so the
 * photometric term sees gradients from any viewpoint.  The renderer
 * ray-casts a pinhole camera (fx, fy, cx, cy; x right, y down, z forward)
 * and produces, per pixel:
 *   depth  uint16, 1000 units/m, Gaussian noise sigma = noise_mm * z^2 (z in m),
 *          a seeded fraction of dropout holes, 0 outside [0.5 m, 5 m];
 *   rgb    3 bytes;
 *   vert4  camera-frame vertex (x, y, z, 1), z = 0 where nothing valid is hit
 *          (the RGBA32F layout of a predicted model vertex map);
 *   norm4  camera-frame normal (nx, ny, nz, 1);
 *   rgba4  4 bytes (the RGBA8 layout of a predicted model image).
 * The noise-free outputs stand in for the surfel model's splatted prediction.
 *
 * Perlin noise follows K. Perlin, "Improving Noise", SIGGRAPH 2002 (fade
 * 6t^5-15t^4+10t^3, 12 gradient directions, seeded permutation); random
 * numbers come from splitmix64 and Box-Muller.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define SYNTH_MAX_BOXES 64

typedef struct {
    float lo[3], hi[3];
    float albedo[3];
} synth_box;

typedef struct {
    float room_lo[3], room_hi[3];
    float wall_albedo[6][3];
    int nbox;
    synth_box box[SYNTH_MAX_BOXES];
    float tex_freq;
    int perm[512];
} synth_scene;

static uint64_t splitmix64(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static double urand(uint64_t *s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

static double nrand(uint64_t *s)
{
    double u1 = urand(s), u2 = urand(s);
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

int synth_scene_size(void) { return (int)sizeof(synth_scene); }

/* Room [-2.5,2.5] x [-1.4,1.4] x [-2.5,2.5] (y down), nbox boxes kept away
 * from the camera region around (0, 0, -1.5). */
void synth_scene_create(uint64_t seed, int nbox, synth_scene *s)
{
    uint64_t st = seed * 0x2545F4914F6CDD1Dull + 12345u;
    memset(s, 0, sizeof *s);
    const float lo[3] = {-2.5f, -1.4f, -2.5f}, hi[3] = {2.5f, 1.4f, 2.5f};
    memcpy(s->room_lo, lo, sizeof lo);
    memcpy(s->room_hi, hi, sizeof hi);
    for (int i = 0; i < 6; ++i)
        for (int c = 0; c < 3; ++c) s->wall_albedo[i][c] = (float)(0.45 + 0.55 * urand(&st));
    if (nbox > SYNTH_MAX_BOXES) nbox = SYNTH_MAX_BOXES;
    int n = 0, tries = 0;
    while (n < nbox && tries < 10000) {
        ++tries;
        synth_box b;
        float sz[3], c[3];
        for (int k = 0; k < 3; ++k) sz[k] = (float)(0.15 + 0.65 * urand(&st));
        c[0] = (float)(-2.2 + 4.4 * urand(&st));
        c[2] = (float)(-0.6 + 2.9 * urand(&st));
        /* half on the floor, half floating */
        if (urand(&st) < 0.5) c[1] = hi[1] - 0.5f * sz[1];
        else c[1] = (float)(-1.0 + 2.0 * urand(&st));
        for (int k = 0; k < 3; ++k) {
            b.lo[k] = c[k] - 0.5f * sz[k];
            b.hi[k] = c[k] + 0.5f * sz[k];
            if (b.lo[k] < lo[k]) b.lo[k] = lo[k];
            if (b.hi[k] > hi[k]) b.hi[k] = hi[k];
        }
        /* keep a 1.2 m clearance box around the camera region */
        if (b.hi[0] > -1.2f && b.lo[0] < 1.2f && b.hi[1] > -0.9f && b.lo[1] < 0.9f &&
            b.hi[2] > -2.5f && b.lo[2] < -0.3f)
            continue;
        for (int k = 0; k < 3; ++k) b.albedo[k] = (float)(0.3 + 0.7 * urand(&st));
        s->box[n++] = b;
    }
    s->nbox = n;
    s->tex_freq = 6.0f;
    int p[256];
    for (int i = 0; i < 256; ++i) p[i] = i;
    for (int i = 255; i > 0; --i) {
        int j = (int)(splitmix64(&st) % (uint64_t)(i + 1));
        int t = p[i]; p[i] = p[j]; p[j] = t;
    }
    for (int i = 0; i < 512; ++i) s->perm[i] = p[i & 255];
}

static inline double fade(double t) { return t * t * t * (t * (t * 6 - 15) + 10); }
static inline double lerp(double t, double a, double b) { return a + t * (b - a); }
static inline double grad(int hash, double x, double y, double z)
{
    const int h = hash & 15;
    const double u = h < 8 ? x : y;
    const double v = h < 4 ? y : (h == 12 || h == 14 ? x : z);
    return ((h & 1) == 0 ? u : -u) + ((h & 2) == 0 ? v : -v);
}

static double perlin(const int *P, double x, double y, double z)
{
    const double fx = floor(x), fy = floor(y), fz = floor(z);
    const int X = (int)fx & 255, Y = (int)fy & 255, Z = (int)fz & 255;
    x -= fx; y -= fy; z -= fz;
    const double u = fade(x), v = fade(y), w = fade(z);
    const int A = P[X] + Y, AA = P[A] + Z, AB = P[A + 1] + Z;
    const int B = P[X + 1] + Y, BA = P[B] + Z, BB = P[B + 1] + Z;
    return lerp(w, lerp(v, lerp(u, grad(P[AA], x, y, z), grad(P[BA], x - 1, y, z)),
                        lerp(u, grad(P[AB], x, y - 1, z), grad(P[BB], x - 1, y - 1, z))),
                lerp(v, lerp(u, grad(P[AA + 1], x, y, z - 1), grad(P[BA + 1], x - 1, y, z - 1)),
                        lerp(u, grad(P[AB + 1], x, y - 1, z - 1), grad(P[BB + 1], x - 1, y - 1, z - 1))));
}

/* texture value in [0, 1]: two octaves, contrast-sharpened */
static double texture(const synth_scene *s, const double p[3])
{
    const double f = s->tex_freq;
    double n = perlin(s->perm, p[0] * f, p[1] * f, p[2] * f);
    n += 0.5 * perlin(s->perm, p[0] * f * 2.3 + 17.1, p[1] * f * 2.3 + 3.7, p[2] * f * 2.3 + 9.2);
    double t = 0.5 + 1.8 * n;
    if (t < 0) t = 0;
    if (t > 1) t = 1;
    return t;
}

/* nearest hit along o + t d, t > 0; returns t, writes world normal and albedo */
static double cast(const synth_scene *s, const double o[3], const double d[3], double n[3], double alb[3])
{
    double best = INFINITY;
    /* room interior: exit distance per slab */
    for (int k = 0; k < 3; ++k) {
        if (d[k] == 0.0) continue;
        const double b = d[k] > 0 ? s->room_hi[k] : s->room_lo[k];
        const double t = (b - o[k]) / d[k];
        if (t > 0 && t < best) {
            best = t;
            n[0] = n[1] = n[2] = 0;
            n[k] = d[k] > 0 ? -1.0 : 1.0;
            const int wi = 2 * k + (d[k] > 0);
            for (int c = 0; c < 3; ++c) alb[c] = s->wall_albedo[wi][c];
        }
    }
    for (int i = 0; i < s->nbox; ++i) {
        const synth_box *b = &s->box[i];
        double tmin = -INFINITY, tmax = INFINITY;
        int axis = -1;
        int ok = 1;
        for (int k = 0; k < 3 && ok; ++k) {
            if (d[k] == 0.0) {
                if (o[k] < b->lo[k] || o[k] > b->hi[k]) ok = 0;
                continue;
            }
            double t1 = (b->lo[k] - o[k]) / d[k], t2 = (b->hi[k] - o[k]) / d[k];
            if (t1 > t2) { double t = t1; t1 = t2; t2 = t; }
            if (t1 > tmin) { tmin = t1; axis = k; }
            if (t2 < tmax) tmax = t2;
            if (tmin > tmax) ok = 0;
        }
        if (!ok || axis < 0 || tmin <= 0 || tmin >= best) continue;
        best = tmin;
        n[0] = n[1] = n[2] = 0;
        n[axis] = d[axis] > 0 ? -1.0 : 1.0;
        for (int c = 0; c < 3; ++c) alb[c] = b->albedo[c];
    }
    return best;
}

/* T: row-major 4x4 camera-to-world.  Any output pointer may be NULL. */
void synth_render(const synth_scene *s, const double *T, int w, int h, double fx, double fy, double cx,
                  double cy, uint16_t *depth, uint8_t *rgb, float *vert4, float *norm4, uint8_t *rgba4,
                  double noise_mm, double dropout, uint64_t noise_seed)
{
#pragma omp parallel for schedule(dynamic, 8)
    for (int v = 0; v < h; ++v) {
        uint64_t st = noise_seed * 0x9E3779B97F4A7C15ull + (uint64_t)v * 0xD1B54A32D192ED03ull + 7u;
        for (int u = 0; u < w; ++u) {
            const size_t i = (size_t)v * w + u;
            const double dc[3] = {(u - cx) / fx, (v - cy) / fy, 1.0};
            double d[3], o[3] = {T[3], T[7], T[11]};
            for (int r = 0; r < 3; ++r) d[r] = T[4 * r] * dc[0] + T[4 * r + 1] * dc[1] + T[4 * r + 2] * dc[2];
            double nw[3] = {0, 0, 0}, alb[3] = {0, 0, 0};
            const double t = cast(s, o, d, nw, alb);
            const int hit = isfinite(t) && t >= 0.5 && t <= 5.0;
            double col[3] = {0, 0, 0};
            if (hit) {
                const double p[3] = {o[0] + t * d[0], o[1] + t * d[1], o[2] + t * d[2]};
                const double tx = texture(s, p);
                for (int c = 0; c < 3; ++c) col[c] = 255.0 * alb[c] * (0.2 + 0.8 * tx);
            }
            /* noise draws happen for every pixel so streams do not depend on hits */
            const double g = nrand(&st);
            const double drop = urand(&st);
            if (depth) {
                uint16_t z = 0;
                if (hit) {
                    double mm = t * 1000.0 + g * noise_mm * t * t;
                    mm = floor(mm + 0.5);
                    if (mm >= 500.0 && mm <= 5000.0 && drop >= dropout) z = (uint16_t)mm;
                }
                depth[i] = z;
            }
            if (rgb)
                for (int c = 0; c < 3; ++c) rgb[3 * i + c] = (uint8_t)(col[c] > 255 ? 255 : col[c]);
            if (rgba4) {
                for (int c = 0; c < 3; ++c) rgba4[4 * i + c] = (uint8_t)(col[c] > 255 ? 255 : col[c]);
                rgba4[4 * i + 3] = 255;
            }
            if (vert4 || norm4) {
                float vv[4] = {0, 0, 0, 0}, nn[4] = {0, 0, 0, 0};
                if (hit) {
                    vv[0] = (float)(t * dc[0]); vv[1] = (float)(t * dc[1]); vv[2] = (float)t; vv[3] = 1.0f;
                    /* camera-frame normal R^T n */
                    for (int r = 0; r < 3; ++r)
                        nn[r] = (float)(T[r] * nw[0] + T[4 + r] * nw[1] + T[8 + r] * nw[2]);
                    nn[3] = 1.0f;
                }
                if (vert4) memcpy(vert4 + 4 * i, vv, sizeof vv);
                if (norm4) memcpy(norm4 + 4 * i, nn, sizeof nn);
            }
        }
    }
}

/*
 * opiv_synth.c -- seeded synthetic PIV frame pairs.  See opiv_synth.h.
 *
 * Random numbers are counter based (SplitMix64 of seed, stream and index),
 * so each particle and each pixel's noise is a pure function of its index;
 * threads only split the rows to render.
 */
#define _GNU_SOURCE
#include "opiv_synth.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define MARGIN 16.0   /* particles are seeded this far outside the frame */
#define RADIUS 4      /* rendering half-extent in pixels */

static inline uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

static inline double u01(uint64_t seed, uint64_t stream, uint64_t ctr)
{
    const uint64_t h = splitmix64(seed ^ splitmix64((stream << 48) ^ ctr));
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

void opiv_synth_field(const opiv_synth_desc* d, double x, double y, double* dx, double* dy)
{
    switch (d->field) {
    case OPIV_FIELD_ROTATION: {
        const double cx = 0.5 * (d->w - 1), cy = 0.5 * (d->h - 1);
        const double omega = d->p1 > 0 ? d->p0 / d->p1 : 0.0;
        *dx = -omega * (y - cy);
        *dy = omega * (x - cx);
        break;
    }
    case OPIV_FIELD_LAMB_OSEEN: {
        /* v(r) = A (1 - exp(-r^2/rc^2)) / r, peak at r = 1.12091 rc where
         * (1 - exp(-1.25644)) / 1.12091 = 0.638161 / rc */
        const double rx = x - d->p0, ry = y - d->p1, rc = d->p2;
        const double r2 = rx * rx + ry * ry;
        if (r2 < 1e-12 || rc <= 0) { *dx = 0; *dy = 0; break; }
        const double a = d->p3 * rc / 0.6381726863389515;
        const double vt_over_r = a * (1.0 - exp(-r2 / (rc * rc))) / r2;
        *dx = -vt_over_r * ry;
        *dy = vt_over_r * rx;
        break;
    }
    default:
        *dx = d->p0;
        *dy = d->p1;
        break;
    }
}

typedef struct { double x, y, i0; } particle;

typedef struct {
    const opiv_synth_desc* d;
    const particle* parts;     /* sorted by row bucket */
    const size_t* bucket;      /* start index per row bucket, nb+1 entries */
    long nb;                   /* number of buckets (rows of the extended domain) */
    int frame;                 /* 0 = A, 1 = B (noise stream) */
    void* out;
    uint32_t row0, row1;
} render_arg;

static void* render_rows(void* p)
{
    render_arg* a = (render_arg*)p;
    const opiv_synth_desc* d = a->d;
    const uint32_t W = d->w;
    const double k = 8.0 / (d->diameter * d->diameter);
    double* acc = (double*)malloc(sizeof(double) * W);
    for (uint32_t y = a->row0; y < a->row1; ++y) {
        memset(acc, 0, sizeof(double) * W);
        /* particles whose bucket row lies within RADIUS+1 of y */
        long b0 = (long)y + (long)MARGIN - RADIUS - 1, b1 = (long)y + (long)MARGIN + RADIUS + 1;
        if (b0 < 0) b0 = 0;
        if (b1 > a->nb - 1) b1 = a->nb - 1;
        for (size_t i = a->bucket[b0]; i < a->bucket[b1 + 1]; ++i) {
            const particle* q = &a->parts[i];
            const double ddy = (double)y - q->y;
            if (fabs(ddy) > RADIUS) continue;
            const double ey = ddy * ddy;
            long x0 = (long)floor(q->x) - RADIUS, x1 = (long)floor(q->x) + RADIUS + 1;
            if (x0 < 0) x0 = 0;
            if (x1 > (long)W - 1) x1 = (long)W - 1;
            for (long x = x0; x <= x1; ++x) {
                const double ddx = (double)x - q->x;
                acc[x] += q->i0 * exp(-k * (ddx * ddx + ey));
            }
        }
        for (uint32_t x = 0; x < W; ++x) {
            double v = acc[x];
            if (d->noise_sigma > 0) {
                const uint64_t pix = (uint64_t)y * W + x;
                const double u1 = 1.0 - u01(d->seed, 2 + a->frame, 2 * pix);
                const double u2 = u01(d->seed, 2 + a->frame, 2 * pix + 1);
                v += d->noise_sigma * sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
            }
            const size_t o = (size_t)y * W + x;
            switch (d->dtype) {
            case OPIV_SYNTH_U8: {
                double r = floor(v + 0.5);
                r = r < 0 ? 0 : (r > 255 ? 255 : r);
                ((uint8_t*)a->out)[o] = (uint8_t)r;
                break;
            }
            case OPIV_SYNTH_U16: {
                double r = floor(v + 0.5);
                r = r < 0 ? 0 : (r > 65535 ? 65535 : r);
                ((uint16_t*)a->out)[o] = (uint16_t)r;
                break;
            }
            case OPIV_SYNTH_F32: ((float*)a->out)[o] = (float)v; break;
            default: ((double*)a->out)[o] = v; break;
            }
        }
    }
    free(acc);
    return NULL;
}

static void render(const opiv_synth_desc* d, particle* parts, size_t n, int frame, void* out)
{
    /* bucket by floor(y) over the extended domain [-MARGIN-RADIUS, h+MARGIN+RADIUS) */
    const long nb = (long)d->h + 2 * (long)MARGIN + 2;
    size_t* count = (size_t*)calloc((size_t)nb + 1, sizeof(size_t));
    long* row = (long*)malloc(sizeof(long) * n);
    for (size_t i = 0; i < n; ++i) {
        long r = (long)floor(parts[i].y) + (long)MARGIN;
        if (r < 0) r = 0;
        if (r > nb - 1) r = nb - 1;
        row[i] = r;
        count[r + 1]++;
    }
    for (long r = 0; r < nb; ++r) count[r + 1] += count[r];
    particle* sorted = (particle*)malloc(sizeof(particle) * (n ? n : 1));
    size_t* fill = (size_t*)malloc(sizeof(size_t) * (size_t)nb);
    memcpy(fill, count, sizeof(size_t) * (size_t)nb);
    for (size_t i = 0; i < n; ++i) sorted[fill[row[i]]++] = parts[i];   /* stable */

    int nt = d->nthreads > 0 ? d->nthreads : 8;
    if (nt > 64) nt = 64;
    pthread_t th[64];
    render_arg args[64];
    const uint32_t band = (d->h + nt - 1) / nt;
    int used = 0;
    for (int t = 0; t < nt; ++t) {
        const uint32_t r0 = t * band, r1 = (t + 1) * band > d->h ? d->h : (t + 1) * band;
        if (r0 >= r1) break;
        args[t] = (render_arg){ d, sorted, count, nb, frame, out, r0, r1 };
        pthread_create(&th[t], NULL, render_rows, &args[t]);
        ++used;
    }
    for (int t = 0; t < used; ++t) pthread_join(th[t], NULL);
    free(fill);
    free(sorted);
    free(row);
    free(count);
}

int opiv_synth_pair(const opiv_synth_desc* d, void* frame_a, void* frame_b)
{
    if (!d || d->w == 0 || d->h == 0 || d->diameter <= 0) return -1;
    const double ew = d->w + 2 * MARGIN, eh = d->h + 2 * MARGIN;
    const size_t n = (size_t)llround(d->density * ew * eh);
    particle* pa = (particle*)malloc(sizeof(particle) * (n ? n : 1));
    particle* pb = (particle*)malloc(sizeof(particle) * (n ? n : 1));
    for (size_t i = 0; i < n; ++i) {
        pa[i].x = -MARGIN + ew * u01(d->seed, 1, 3 * i);
        pa[i].y = -MARGIN + eh * u01(d->seed, 1, 3 * i + 1);
        pa[i].i0 = d->i0_lo + (d->i0_hi - d->i0_lo) * u01(d->seed, 1, 3 * i + 2);
        double dx, dy;
        opiv_synth_field(d, pa[i].x, pa[i].y, &dx, &dy);
        pb[i].x = pa[i].x + dx;
        pb[i].y = pa[i].y + dy;
        pb[i].i0 = pa[i].i0;
    }
    render(d, pa, n, 0, frame_a);
    render(d, pb, n, 1, frame_b);
    free(pb);
    free(pa);
    return 0;
}

/*
 * opiv_synth.h -- seeded synthetic PIV frame pairs (particle images).
 *
 * Frame A: Gaussian particles I0*exp(-8 r^2 / d^2) at seeded uniform
 * positions on a zero background, plus N(0, sigma) noise.  Frame B: the same
 * particles moved by a displacement field evaluated at each particle, re-
 * rendered, with fresh noise.  Everything is driven by counter-based
 * SplitMix64 streams, so a (descriptor, seed) pair always renders the same
 * bytes regardless of thread count.
 */
#ifndef OPIV_SYNTH_H
#define OPIV_SYNTH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OPIV_SYNTH_U8 = 0, OPIV_SYNTH_U16 = 1, OPIV_SYNTH_F32 = 2, OPIV_SYNTH_F64 = 3 };

enum {
    OPIV_FIELD_UNIFORM = 0,     /* p0 = dx, p1 = dy */
    OPIV_FIELD_ROTATION = 1,    /* solid body about the image centre; p0 = |d| at radius p1 */
    OPIV_FIELD_LAMB_OSEEN = 2   /* p0,p1 = centre, p2 = core radius, p3 = peak |d| */
};

typedef struct {
    uint32_t w, h;
    uint64_t seed;
    double density;      /* particles per pixel */
    double diameter;     /* particle diameter d (px) */
    double i0_lo, i0_hi; /* peak intensity ~ U(lo, hi) */
    double noise_sigma;  /* additive Gaussian noise */
    int field;
    double p0, p1, p2, p3;
    int dtype;           /* OPIV_SYNTH_* ; integer types are rounded and clamped */
    int nthreads;
} opiv_synth_desc;

/* Renders frames a and b (w*h elements each, row-major).  Returns 0 or -1. */
int opiv_synth_pair(const opiv_synth_desc* d, void* frame_a, void* frame_b);

/* displacement of the field at (x, y) */
void opiv_synth_field(const opiv_synth_desc* d, double x, double y, double* dx, double* dy);

#ifdef __cplusplus
}
#endif
#endif

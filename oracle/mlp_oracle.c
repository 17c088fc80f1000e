/*
 * CPU oracle of the learned work predictor (6-32-32-1 MLP).  TEST
 * INFRASTRUCTURE ONLY: the GPU feature, forward and backward kernels are
 * checked against these functions.  The predictor has no counterpart in the
 * reference; it only changes the order in which pairs are packed.
 *
 * Parameter vector (1,313 floats):
 *   W1[32][6], b1[32], W2[32][32], b2[32], W3[32], b3
 * Forward, with round16() = round to IEEE binary16 (nearest even) when
 * `fp16` is set (the GPU's storage format), identity otherwise:
 *   x  = round16((raw - mean) * inv_std)
 *   h1 = round16(relu(W1 x + b1)),  h2 = round16(relu(W2 h1 + b2))
 *   y  = W3 h2 + b3            (predicted log(1 + cells))
 * with weights also passed through round16(); sums accumulate in float.
 * Loss of a minibatch: mean over pairs of 0.5 * (y - log1p(cells))^2.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#define NIN 6
#define NH 32
#define NPARAM (NH * NIN + NH + NH * NH + NH + NH + 1)

/* float -> binary16 -> float, round to nearest even (no denormal flush) */
static float round16(float f) {
    uint32_t x;
    memcpy(&x, &f, 4);
    uint32_t sign = x & 0x80000000u;
    uint32_t ax = x & 0x7FFFFFFFu;
    if (ax >= 0x7F800000u) return f; /* inf / nan */
    if (ax >= 0x477FF000u) {         /* rounds to >= 65520: overflow to inf */
        uint32_t r = sign | 0x7F800000u;
        float o;
        memcpy(&o, &r, 4);
        return o;
    }
    float a;
    memcpy(&a, &ax, 4);
    float out;
    if (ax < 0x38800000u) { /* below 2^-14: half subnormal, quantum 2^-24 */
        float q = a * 16777216.0f;          /* exact scaling by 2^24 */
        float r = nearbyintf(q);            /* ties to even under default rounding */
        out = r / 16777216.0f;
    } else {
        uint32_t lsb = (ax >> 13) & 1u;
        uint32_t rnd = ax + 0x0FFFu + lsb;
        rnd &= 0xFFFFE000u;
        memcpy(&out, &rnd, 4);
    }
    uint32_t o;
    memcpy(&o, &out, 4);
    o |= sign;
    memcpy(&out, &o, 4);
    return out;
}

float ora_round16(float f) { return round16(f); }

static float rq(float v, int fp16) { return fp16 ? round16(v) : v; }

/* Forward of one pair; returns y and fills the intermediate vectors. */
static float forward1(const float *p, const float *raw, const float *norm, int fp16, float *x,
                      float *z1, float *h1, float *z2, float *h2) {
    const float *W1 = p, *b1 = p + NH * NIN, *W2 = b1 + NH, *b2 = W2 + NH * NH, *W3 = b2 + NH,
                *b3 = W3 + NH;
    for (int k = 0; k < NIN; ++k) x[k] = rq((raw[k] - norm[k]) * norm[NIN + k], fp16);
    for (int i = 0; i < NH; ++i) {
        float a = rq(b1[i], fp16);
        for (int k = 0; k < NIN; ++k) a += rq(W1[i * NIN + k], fp16) * x[k];
        z1[i] = a;
        h1[i] = rq(a > 0.f ? a : 0.f, fp16);
    }
    for (int i = 0; i < NH; ++i) {
        float a = rq(b2[i], fp16);
        for (int k = 0; k < NH; ++k) a += rq(W2[i * NH + k], fp16) * h1[k];
        z2[i] = a;
        h2[i] = rq(a > 0.f ? a : 0.f, fp16);
    }
    float y = rq(b3[0], fp16);
    for (int k = 0; k < NH; ++k) y += rq(W3[k], fp16) * h2[k];
    return y;
}

int ora_mlp_nparam(void) { return NPARAM; }

/* y[n] for raw features feat[n][6] */
void ora_mlp_forward(const float *params, const float *norm, const float *feat, int64_t n,
                     int fp16, float *y) {
    float x[NIN], z1[NH], h1[NH], z2[NH], h2[NH];
    for (int64_t q = 0; q < n; ++q)
        y[q] = forward1(params, feat + q * NIN, norm, fp16, x, z1, h1, z2, h2);
}

/* Mean loss and its gradient (double accumulation) over pairs [0, n). */
double ora_mlp_grad(const float *params, const float *norm, const float *feat,
                    const float *target, int64_t n, int fp16, double *grad) {
    const float *W2 = params + NH * NIN + NH, *W3 = W2 + NH * NH + NH;
    double *gW1 = grad, *gb1 = grad + NH * NIN, *gW2 = gb1 + NH, *gb2 = gW2 + NH * NH,
           *gW3 = gb2 + NH, *gb3 = gW3 + NH;
    for (int k = 0; k < NPARAM; ++k) grad[k] = 0.0;
    double loss = 0.0;
    float x[NIN], z1[NH], h1[NH], z2[NH], h2[NH];
    for (int64_t q = 0; q < n; ++q) {
        float y = forward1(params, feat + q * NIN, norm, fp16, x, z1, h1, z2, h2);
        double d = (double)y - (double)target[q];
        loss += 0.5 * d * d;
        double d3 = d / (double)n;
        gb3[0] += d3;
        double d2[NH];
        for (int j = 0; j < NH; ++j) {
            gW3[j] += d3 * h2[j];
            d2[j] = z2[j] > 0.f ? d3 * (double)rq(W3[j], fp16) : 0.0;
        }
        for (int j = 0; j < NH; ++j) {
            gb2[j] += d2[j];
            for (int k = 0; k < NH; ++k) gW2[j * NH + k] += d2[j] * h1[k];
        }
        for (int k = 0; k < NH; ++k) {
            double a = 0.0;
            for (int j = 0; j < NH; ++j) a += (double)rq(W2[j * NH + k], fp16) * d2[j];
            double d1 = z1[k] > 0.f ? a : 0.0;
            gb1[k] += d1;
            for (int m = 0; m < NIN; ++m) gW1[k * NIN + m] += d1 * x[m];
        }
    }
    return n > 0 ? loss / (double)n : 0.0;
}

/* Raw features of one extension pair from u8 codes:
 * qlen, tlen, h0, w, mismatches among the first lim = min(32, qlen, tlen)
 * positions of the best diagonal within three of the main one, query GC
 * fraction.  Diagonal d >= 0 compares q[j] with t[j + d], d < 0 compares
 * q[j - d] with t[j], j < lim, and is taken only when both windows lie inside
 * their sequences (an indel near the start leaves few mismatches on a
 * neighbouring diagonal, where the main one looks like a random target). */
void ora_pair_features(const uint8_t *q, int32_t qlen, const uint8_t *t, int32_t tlen, int32_t h0,
                       int32_t w, float *out) {
    int lim = qlen < tlen ? qlen : tlen;
    if (lim > 32) lim = 32;
    int mm = 1 << 30;
    for (int d = -3; d <= 3; ++d) {
        const int qo = d < 0 ? -d : 0, to = d > 0 ? d : 0;
        if (qo + lim > qlen || to + lim > tlen) continue;
        int c = 0;
        for (int j = 0; j < lim; ++j) c += q[qo + j] != t[to + j];
        if (c < mm) mm = c;
    }
    int gc = 0;
    for (int j = 0; j < qlen; ++j) gc += q[j] == 1 || q[j] == 2;
    out[0] = (float)qlen;
    out[1] = (float)tlen;
    out[2] = (float)h0;
    out[3] = (float)w;
    out[4] = (float)mm;
    out[5] = qlen > 0 ? (float)gc / (float)qlen : 0.f;
}

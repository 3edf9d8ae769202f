/*
 * orc_exp.h — e^x shared by the oracle's restatements (TEST
 * INFRASTRUCTURE ONLY).  Include after `real` and ORC_DOUBLE are defined.
 *
 * The FP64 build calls libm's exp, as the reference does.  The FP32 build
 * uses Cody-Waite reduction x = k ln2 + r (ln2 split into a 17-bit head and a
 * tail), a degree-7 Taylor polynomial of e^r by Horner's rule and exact
 * scaling by 2^k: only basic IEEE operations, so the GPU's version of the
 * same steps gives the same float.
 */
#ifndef ORC_EXP_H
#define ORC_EXP_H

#if ORC_DOUBLE
static inline double orc_exp(double x) { return exp(x); }
#else
static inline float orc_exp(float x)
{
    if (x > 80.0f) x = 80.0f;
    if (x < -80.0f) x = -80.0f;
    float k = floorf((x * 1.44269504088896341f) + 0.5f);
    float r = (x - (k * 0.693145751953125f)) - (k * 1.42860682030941723e-6f);
    float p = 1.98412698412698413e-4f;
    p = (p * r) + 1.38888888888888889e-3f;
    p = (p * r) + 8.33333333333333333e-3f;
    p = (p * r) + 4.16666666666666667e-2f;
    p = (p * r) + 1.66666666666666667e-1f;
    p = (p * r) + 0.5f;
    p = (p * r) + 1.0f;
    p = (p * r) + 1.0f;
    return ldexpf(p, (int)k);
}
#endif

#endif /* ORC_EXP_H */

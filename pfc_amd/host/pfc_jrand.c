/* pfc_jrand.c -- java.util.Random's specified algorithm (see pfc_jrand.h). */
#include "pfc_jrand.h"

#define JR_MUL 0x5DEECE66DULL
#define JR_ADD 0xBULL
#define JR_MASK ((1ULL << 48) - 1)

void pfc_jrand_seed(pfc_jrand* r, int64_t seed) { r->s = ((uint64_t)seed ^ JR_MUL) & JR_MASK; }

int32_t pfc_jrand_next(pfc_jrand* r, int bits) {
    r->s = (r->s * JR_MUL + JR_ADD) & JR_MASK;
    /* (int)(seed >>> (48 - bits)): the low 32 bits of the shifted state */
    return (int32_t)(uint32_t)(r->s >> (48 - bits));
}

int32_t pfc_jrand_next_int(pfc_jrand* r) { return pfc_jrand_next(r, 32); }

double pfc_jrand_next_double(pfc_jrand* r) {
    const int64_t hi = (int64_t)pfc_jrand_next(r, 26);
    const int64_t lo = (int64_t)pfc_jrand_next(r, 27);
    return (double)((hi << 27) + lo) * (1.0 / 9007199254740992.0);   /* 2^-53 */
}

void pfc_jrand_doubles(int64_t seed, double* out, long n) {
    pfc_jrand r;
    pfc_jrand_seed(&r, seed);
    for (long i = 0; i < n; i++) out[i] = pfc_jrand_next_double(&r);
}

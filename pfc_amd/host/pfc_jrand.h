/*
 * pfc_jrand.h -- java.util.Random, the generator pfc-init draws its noise
 * from (src/PFCInit.java:294, 320).  The algorithm is the one the Java
 * platform specifies for the class: a 48-bit linear congruential generator
 * (multiplier 0x5DEECE66D, increment 0xB), the seed scrambled by XOR with the
 * multiplier, next(bits) taking the top bits, and nextDouble built from 26 +
 * 27 bits.  With the same seed it yields the same numbers as the JVM.
 */
#ifndef PFC_JRAND_H
#define PFC_JRAND_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint64_t s;   /* 48-bit state */
} pfc_jrand;

void pfc_jrand_seed(pfc_jrand* r, int64_t seed);
int32_t pfc_jrand_next(pfc_jrand* r, int bits);
int32_t pfc_jrand_next_int(pfc_jrand* r);
double pfc_jrand_next_double(pfc_jrand* r);

/* n successive nextDouble() values after seeding with seed */
void pfc_jrand_doubles(int64_t seed, double* out, long n);

#ifdef __cplusplus
}
#endif

#endif

/* jms_sequence.h -- frozen-phonon variant sequencing on the host.
 * The draw is that of CJMultiSlice::GetRandomVariantSequence
 * (JMultiSlice.cpp:1571-1676) over a CRng stream (rng.cpp:30-89).
 */
#ifndef JMS_SEQUENCE_H
#define JMS_SEQUENCE_H

#ifdef __cplusplus
extern "C" {
#endif

/* nseq consecutive variant sequences from one generator seeded with `seed`:
 * out[k*nobjslc + j] is the variant of object slice j in sequence k
 * (structure slice objslc[j], nvarslc[objslc[j]] variants).  This is the
 * serial order in which a scan draws them: position-major, pass-minor. */
int jms_variant_sequences(int seed, int nscslc, const int* nvarslc, int nobjslc, const int* objslc,
                          int nseq, int* out);

#ifdef __cplusplus
}
#endif
#endif

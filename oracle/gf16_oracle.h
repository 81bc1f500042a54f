/*
 * ORACLE — test infrastructure, not product code.
 *
 * A plain-C restatement of ParPar's GF(2^16) recovery path (PAR2ProcCPU's
 * results), used only by tests/, the smoke check and the benchmark's CPU
 * baseline leg, and only as the checker.  Each function names the reference
 * file:line it restates.  Pinning (what proves each function right) is listed
 * in gf16_oracle.c's header; every function here is pinned.
 */
#ifndef GF16_ORACLE_H
#define GF16_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

void po_init(void);

/* GF(2^16) arithmetic, polynomial 0x1100B (gf16/gf16_global.h:9-10) */
uint16_t po_mul_slow(uint16_t a, uint16_t b);
uint16_t po_mul(uint16_t a, uint16_t b);
uint16_t po_exp(uint32_t v);
uint16_t po_log(uint16_t x);

/* PAR2 coefficients (gf16/gfmat_coeff.c:6-77) */
uint16_t po_input_log(uint16_t input);
uint16_t po_coeff_log(uint16_t inputLog, uint16_t exp);
uint16_t po_coeff(uint16_t input, uint16_t exp);

/* gf16_cksum (gf16/gf16_cksum_base.h:13-59, gf16/gf16_checksum_generic.h) */
void po_cksum(const uint8_t* src, size_t srcLen, size_t sliceLen, unsigned width, uint8_t* ck);
int po_cksum_verify(const uint8_t* data, size_t len, const uint8_t* ck, unsigned width);

/* dst ^= coeff * src over ceil(len/2) little-endian words (odd tail = low byte) */
void po_muladd_region(uint8_t* dst, const uint8_t* src, size_t len, uint16_t coeff);

/* Whole recovery as PAR2ProcCPU computes it (gf16/controller_cpu.cpp:463-537):
 * outputs[r] = sum_i c(r,i) * pad(inputs[i]) for r < nOut; coeffs (optional)
 * is coeffs[i*nOut + r], otherwise c(r,i) = po_coeff(inputNums[i], exps[r]).
 * threads <= 0 uses every OpenMP thread.  Returns the thread count used. */
int po_recover(unsigned nIn, const uint8_t* inputs, size_t inStride, const size_t* inLens,
               const uint16_t* inputNums, const uint16_t* coeffs, unsigned nOut,
               const uint16_t* exps, size_t sliceSize, uint8_t* outputs, int threads);

/* Synthetic inputs (oracle/po_prng.h) */
void po_prng_fill_c(uint64_t seed, uint32_t job, uint32_t slice, uint64_t offset, uint8_t* dst, size_t len);
/* gf16_cksum of synthetic slices first .. first+count-1 of (seed, job), lens[k]
 * bytes each (NULL: sliceLen) zero-padded to sliceLen: po_cksum of
 * po_prng_fill's bytes, slices in parallel (threads <= 0: all). */
void po_prng_cksums(uint64_t seed, uint32_t job, uint32_t first, unsigned count, const uint32_t* lens,
                    size_t sliceLen, unsigned width, uint8_t* out, int threads);

/* Hashes of the per-slice path (hasher/hasher_base.h:83-198): MD5 (RFC 1321) and
 * CRC-32 (IEEE 802.3, reflected 0xEDB88320) of data zero-padded to padLen. */
void po_md5(const uint8_t* data, size_t len, size_t padLen, uint8_t out[16]);
uint32_t po_crc32(const uint8_t* data, size_t len, size_t padLen);

#ifdef __cplusplus
}
#endif
#endif

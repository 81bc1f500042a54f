/*
 * parpar_hip — MI355X (gfx950) GF(2^16) recovery engine for ParPar.
 *
 * C ABI of libparpar_hip.so.  Plain pointers and sizes only; every entry point
 * returns an int status (PPH_OK = 0, negative = error; pph_strerror() names it)
 * unless stated otherwise, and no C++ exception crosses it.
 *
 * Two layers:
 *
 *   pph_proc_*  one recovery backend, entry for entry the plugin interface a
 *               ParPar backend implements (IPAR2ProcBackend, gf16/controller.h:77-217),
 *               with PAR2ProcOCL's semantics (gf16/controller_ocl.cpp) for staging,
 *               batching and checksums.  Host buffers are caller-owned; an add
 *               returns once the caller's buffer may be reused (the point where
 *               PAR2ProcOCL's add future resolves, controller_ocl.cpp:296).
 *               Calls on one pph_proc are not thread-safe (ParPar drives a
 *               backend from one thread, controller.h:13-35).
 *
 *   pph_*       device-resident building blocks on HBM buffers (slot arrays),
 *               used by the batched and serving drivers and the benchmark:
 *               synthetic fill, input preparation (zero-pad + gf16_cksum +
 *               log-domain conversion), the plain and dense mul-add kernels and
 *               output checksum verification.
 *
 * Device slice format ("slot", one per input or recovery slice):
 *   [slice data, slice_size bytes][zero][gf16_cksum block, cksum_size bytes at
 *   ck_off = round_up(slice_size, 16)][zero to slot_bytes]
 * The reference's OpenCL backend keeps the same [data][cksum] shape
 * (gf16/controller_ocl.cpp:90,285-288); the checksum sits at a 16-byte aligned
 * offset here so that an odd sub-slice never shares a GF(2^16) word with it.
 * Plain-form inputs are stored in log domain after preparation: word x is
 * replaced by log2(x) (x != 0) or 0xFFFF (x == 0).
 */
#ifndef PARPAR_HIP_H
#define PARPAR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define PPH_OK 0
#define PPH_E_INVALID (-1)      /* argument out of range */
#define PPH_E_NOMEM (-2)        /* device or pinned host allocation failed */
#define PPH_E_STATE (-3)        /* call out of sequence */
#define PPH_E_DEVICE (-4)       /* HIP runtime error; pph_last_hip_error() has the code */
#define PPH_E_UNSUPPORTED (-5)  /* no HIP device / unsupported configuration */

/* IPAR2ProcBackend::canAdd results (gf16/controller.h:38-43) */
#define PPH_ADD_OK 0
#define PPH_ADD_OK_BUSY 1
#define PPH_ADD_FULL 2

/* mul-add forms */
#define PPH_FORM_AUTO 0
#define PPH_FORM_PLAIN 1   /* LDS log/exp streaming kernel */
#define PPH_FORM_DENSE 2   /* GF(2) bit-matrix contraction on int8 MFMA */
#define PPH_FORM_BILINEAR 3  /* rank-81 bilinear formula on fp4 MFMA over 0/1 bit planes (pph_muladd_bilinear) */

const char* pph_strerror(int status);
int pph_last_hip_error(void);
int pph_device_count(void);
/* Device name and CU count of a device; name may be NULL. */
int pph_device_info(int device, char* name, size_t name_len, int* cu_count);
/* Device memory in bytes: the memory_global of ParPar's device info
 * (src/gf.cc:777), which its planner sizes device shares by
 * (lib/par2gen.js:509-516). */
int pph_device_memory(int device, size_t* total_bytes);

/* ---- PAR2 coefficient helpers (gf16/gfmat_coeff.c:56-77) ---- */
uint16_t pph_input_log(uint16_t input_num);                     /* gfmat_input_log */
uint16_t pph_coeff_log(uint16_t input_log, uint16_t exponent);  /* gfmat_coeff_log */
uint16_t pph_coeff(uint16_t input_num, uint16_t exponent);      /* gfmat_coeff */
uint16_t pph_gf_mul(uint16_t a, uint16_t b);                    /* multiply in GF(2^16), 0x1100B */

/* Slot geometry for a slice size and checksum width (8, 16, 32 or 64). */
int pph_slot_layout(size_t slice_size, unsigned cksum_size, size_t* ck_off, size_t* slot_bytes);

/* ======================================================================
 * Backend layer: IPAR2ProcBackend (gf16/controller.h:77-217)
 * ====================================================================== */
typedef struct pph_proc pph_proc;

/* Constructor: PAR2ProcOCL(platformId, deviceId, stagingAreas)
 * (gf16/controller_ocl.cpp:27-63).  Fails with PPH_E_UNSUPPORTED without a device. */
int pph_proc_create(int device, unsigned staging_areas, pph_proc** out);
/* PAR2ProcOCL::init (gf16/controller_ocl.cpp:67-90): cksum_size is the checksum
 * width of the CPU method the host selects (Galois16Mul::info().cksumSize,
 * gf16/gf16mul.cpp:215-409; 64 on AVX512+GFNI hosts); input_batch 0 picks the
 * default (256 inputs, fewer where a staging area would pass 512 MiB); form
 * is PPH_FORM_PLAIN (plain kernel), _BILINEAR (the bilinear
 * form, for batches numbered by input and explicit-coefficient batches alike;
 * worth it with large input batches and many recovery slices) or _AUTO (bilinear for input batches of at least 256 with at least
 * 64 recovery slices, plain otherwise; slots prepared by fill_input keep the
 * form chosen then, and a set_recovery_slices that would switch it returns
 * PPH_E_STATE).  setSliceSize must have been called. */
int pph_proc_init(pph_proc* p, unsigned cksum_size, unsigned input_batch, int form);
/* deinit / destructor (controller.h:198-203, controller_ocl.cpp:150-153) */
void pph_proc_destroy(pph_proc* p);
int pph_proc_set_slice_size(pph_proc* p, size_t slice_size);          /* setSliceSize (controller.h:160) */
int pph_proc_set_current_slice_size(pph_proc* p, size_t slice_size);  /* setCurrentSliceSize (:166) */
int pph_proc_set_recovery_slices(pph_proc* p, unsigned n, const uint16_t* exponents); /* (:167) */
int pph_proc_can_add(const pph_proc* p);                              /* canAdd (:168) -> PPH_ADD_* */
int pph_proc_wait_for_add(pph_proc* p);                               /* waitForAdd (:194) */
/* addInput(buffer, size, inputNum, flush) (:169) */
int pph_proc_add_input(pph_proc* p, const void* buffer, size_t size, uint16_t input_num, int flush);
/* addInput(buffer, size, coeffs, flush) (:170): coeffs[r] for every recovery slice r */
int pph_proc_add_input_coeffs(pph_proc* p, const void* buffer, size_t size, const uint16_t* coeffs, int flush);
int pph_proc_dummy_input(pph_proc* p, uint16_t input_num, int flush); /* dummyInput (:171) */
/* fillInput (:172): returns 1 once every staging area is filled, 0 otherwise */
int pph_proc_fill_input(pph_proc* p, const void* buffer);
int pph_proc_flush(pph_proc* p);                                      /* flush (:173) */
/* endInput (:179): flushes and returns when every submitted batch has completed */
int pph_proc_end_input(pph_proc* p);
/* getOutput (:185): copies current_slice_size bytes of recovery slice `index`
 * into output; *valid = gf16_cksum check (copy_cksum_check,
 * gf16/gf16_cksum_base.h:48-59) computed on the host over the bytes as they
 * are copied into output, against the checksum block the mul-add produced. */
int pph_proc_get_output(pph_proc* p, unsigned index, void* output, int* valid);
void pph_proc_discard_output(pph_proc* p);                            /* discardOutput (:186) */
int pph_proc_has_added(const pph_proc* p);                            /* _hasAdded (:189) */
int pph_proc_is_empty(const pph_proc* p);                             /* isEmpty (:182) */
int pph_proc_free_processing_mem(pph_proc* p);                        /* freeProcessingMem (:203) */
unsigned pph_proc_input_batch_size(const pph_proc* p);                /* getInputBatchSize (:205) */
unsigned pph_proc_batches_started(const pph_proc* p);                 /* getBatchesStarted (:208) */
void pph_proc_set_min_input_batch_size(pph_proc* p, unsigned n);      /* setMinInputBatchSize (:212) */
int pph_proc_num_recovery_slices(const pph_proc* p);                  /* getNumRecoverySlices (:157) */
/* Fault injection for tests: XOR `mask` into byte `offset` of recovery slice
 * `index` on the device (the checksum must then report the slice invalid). */
int pph_proc_debug_corrupt_output(pph_proc* p, unsigned index, size_t offset, uint8_t mask);
/* Fault injection for tests: XOR `mask` into byte `offset` of the next
 * recovery slice getOutput copies, in host memory after its device-to-host
 * copy and before the checked copy to the caller (that slice must then be
 * reported invalid while the device copy stays intact). */
int pph_proc_debug_corrupt_host_copy(pph_proc* p, size_t offset, uint8_t mask);
/* Fault injection for tests: the next batch submission fails with `status`
 * (negative).  A failed submission loses its inputs; the backend then keeps
 * that status and returns it from endInput and every getOutput (valid = 0)
 * until a new recovery computation starts (discardOutput, init, a grown
 * slice, freeProcessingMem). */
int pph_proc_debug_fail_next_batch(pph_proc* p, int status);

/* ---- Asynchronous form (the libuv build of the interface, controller.h:79-100,
 * 147-152, 169-185): the caller keeps the staging state on its own thread, as
 * PAR2ProcOCL does (controller_ocl.cpp:303-350,405-426), and names the area
 * and slot of every input; completions come back through these functions,
 * called from the backend's worker thread (they should only post to the
 * caller's loop, as kernel_event_callback does, controller_ocl.cpp:485-494).
 * Staging and submission calls come from one thread, and none overlaps a
 * pending getOutput. */
typedef void (*pph_done_fn)(void* user, int status);
typedef void (*pph_output_fn)(void* user, int status, int valid);
unsigned pph_proc_staging_areas(const pph_proc* p);
/* Copy one input (size <= current slice size, zero padded) into slot `slot`
 * of staging area `area` with its PAR2 input number, or with coeffs[r] for
 * every recovery slice r when coeffs is not NULL (the two addInput forms,
 * controller.h:169-170).  Returns once the caller's buffer may be reused;
 * waits first if the area's previous batch is still running. */
int pph_proc_stage_input(pph_proc* p, unsigned area, unsigned slot, const void* buffer, size_t size,
                         uint16_t input_num, const uint16_t* coeffs);
/* dummyInput (:171): the slot keeps its prepared bytes (pph_proc_fill_slot),
 * only its input number is set. */
int pph_proc_stage_dummy(pph_proc* p, unsigned area, unsigned slot, uint16_t input_num);
/* fillInput (:172) into an explicit slot. */
int pph_proc_fill_slot(pph_proc* p, unsigned area, unsigned slot, const void* buffer);
/* run_kernel (:85): the area's first n_inputs slots as one batch.  first != 0
 * starts a new recovery computation (overwrite; the processingAdd = false of
 * discardOutput, :186-188); otherwise the batch accumulates.  done(user,
 * status) runs exactly once: when the batch's mul-add has completed, or when
 * the submission failed (its status; the backend then keeps that error). */
int pph_proc_submit(pph_proc* p, unsigned area, unsigned n_inputs, int first, pph_done_fn done, void* user);
/* getOutput (:185) run on the worker thread after every task posted before
 * it: done(user, status, valid) once `output` holds the slice. */
int pph_proc_get_output_async(pph_proc* p, unsigned index, void* output, pph_output_fn done, void* user);
/* Wait until every posted completion function has returned. */
int pph_proc_drain(pph_proc* p);

/* ======================================================================
 * Device layer
 * ====================================================================== */
int pph_set_device(int device);
/* Free and total device memory of the current device, bytes. */
int pph_mem_get_info(size_t* free_bytes, size_t* total_bytes);
int pph_malloc(void** ptr, size_t bytes);
int pph_free(void* ptr);
int pph_memcpy_h2d(void* dst, const void* src, size_t bytes);
int pph_memcpy_d2h(void* dst, const void* src, size_t bytes);
int pph_memset(void* dst, int value, size_t bytes);
/* Device-to-device copy by a plain 16-byte streaming kernel (bytes and both
 * pointers 16-byte aligned); the reference point for HBM traffic counters. */
int pph_copy(void* dst, const void* src, size_t bytes, void* stream);
/* Ceiling of the plain form's inner operation, measured: random 16-bit
 * lookups into the 128 KiB exp table in LDS, one 1024-thread workgroup per CU
 * (as the plain kernel runs), 16 lookups per lane per round for `rounds`
 * rounds.  Lookups per second of the timed launch (and its time in ms). */
int pph_lds_gather_ceiling(unsigned rounds, double* lookups_per_s, float* ms);
int pph_synchronize(void);
/* Streams are opaque HIP streams at default priority; NULL = the null stream. */
int pph_stream_create(void** stream);
int pph_stream_destroy(void* stream);
int pph_stream_synchronize(void* stream);
/* Event timer on a stream: start/stop record events, elapsed waits for stop. */
int pph_timer_create(void** timer);
int pph_timer_destroy(void* timer);
int pph_timer_start(void* timer, void* stream);
int pph_timer_stop(void* timer, void* stream);
int pph_timer_elapsed_ms(void* timer, float* ms);

/* Synthetic input: slot k (k < count) of `slots` gets bytes [0, len) of the
 * counter-based stream (seed, job, first_slice + k) and zeros to slot_bytes. */
int pph_fill_random(void* slots, size_t slot_bytes, unsigned count, size_t len,
                    uint64_t seed, uint32_t job, uint32_t first_slice, void* stream);

/* Input preparation, the device half of copy_cksum (gf16/gf16_cksum_base.h:32-46):
 * for slot k, bytes at or past lens[k] (device uint32 array; NULL = slice_size)
 * read as zero, the gf16_cksum of the slice_size-byte slice is stored at ck_off,
 * and, with to_log, every word of the slot is converted to log domain. */
int pph_prep_slots(void* slots, size_t slot_bytes, unsigned count, const uint32_t* lens,
                   size_t slice_size, unsigned cksum_size, int to_log, void* stream);
/* The same, reading data-form slots from `src` (a separate, non-overlapping
 * slot array; NULL = in place) and writing the prepared slots to `dst`. */
int pph_prep_slots_from(void* dst, const void* src, size_t slot_bytes, unsigned count, const uint32_t* lens,
                        size_t slice_size, unsigned cksum_size, int to_log, void* stream);

/* Plain mul-add over log-domain input slots:
 *   out[r] (^)= sum_i c(r,i) * in[i]   for r < n_out, i < n_in, over slot_bytes
 * Coefficients are either c(r,i) = 2^(input_logs[i] * exps[r] mod 65535)
 * (coef_log2 == NULL; input_logs/exps are device uint16 arrays), or explicit:
 * coef_log2[i * coef_stride + r] = 2*log2(c), 0xFFFFFFFF for c == 0.
 * accumulate = 0 overwrites the outputs (first batch), 1 XORs into them.
 * group = outputs per pass (0 = auto).  The work is split evenly over the
 * CUs; output tiles that two workgroups share are XORed together with
 * device-scope atomics, so out_slots must be device memory, and launches on
 * one output array must not overlap. */
int pph_muladd_plain(const void* in_slots, size_t slot_bytes, unsigned n_in,
                     const uint16_t* input_logs, const uint16_t* exps,
                     const uint32_t* coef_log2, unsigned coef_stride,
                     void* out_slots, unsigned n_out, int accumulate, unsigned group, void* stream);

/* Dense mul-add over data-form input slots (prepared with to_log = 0): the same
 * result as pph_muladd_plain with PAR2 coefficients, computed as a GF(2)
 * bit-matrix contraction on int8 MFMA with parity reduction. */
int pph_muladd_dense(const void* in_slots, size_t slot_bytes, unsigned n_in, const uint16_t* input_logs,
                     const uint16_t* exps, void* out_slots, unsigned n_out, int accumulate, void* stream);

/* Bilinear dense form: the same result as pph_muladd_plain / pph_muladd_dense
 * (data-form input slots, PAR2 coefficients), computed from a rank-81 bilinear
 * formula for GF(2^16) multiplication (four levels of Karatsuba): 81 integer
 * GEMMs over 0/1 operands, run as one persistent fp4 MFMA kernel over
 * bit-packed planes (exact: the fp32 sums are integers below 2^24), between
 * plane-building and parity/recombination kernels; 81 MACs per mul-add
 * instead of 256.  Needs a
 * device workspace (pph_bilinear_workspace bytes), allocated on first use,
 * kept for later calls and freed by pph_bilinear_release.  Calls on one device
 * share it: each call's device work starts after the previous call's ends,
 * whatever streams they use. */
int pph_muladd_bilinear(const void* in_slots, size_t slot_bytes, unsigned n_in, const uint16_t* input_logs,
                        const uint16_t* exps, void* out_slots, unsigned n_out, int accumulate, void* stream);
/* The bilinear form with explicit coefficients (decoding, repair): the mul-add
 * that IPAR2ProcBackend::addInput(buf, size, const uint16_t* coeffs, flush)
 * feeds (gf16/controller.h:170; PAR2ProcCPU::set_coeffs' explicit overload,
 * gf16/controller_cpu.cpp:372-385).  coef_log2[i * coef_stride + r] =
 * 2*log2(c), 0xFFFFFFFF for c == 0 (pph_muladd_plain's table), coef_stride >=
 * n_out; data-form input slots; workspace and stream behaviour as
 * pph_muladd_bilinear. */
int pph_muladd_bilinear_coef(const void* in_slots, size_t slot_bytes, unsigned n_in, const uint32_t* coef_log2,
                             unsigned coef_stride, void* out_slots, unsigned n_out, int accumulate, void* stream);
int pph_bilinear_workspace(size_t slot_bytes, unsigned n_in, unsigned n_out, size_t* bytes);
int pph_bilinear_release(void);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The same result as pph_muladd_bilinear, with no
 * hipMalloc, hipFree or other allocation: the GEMM's planes and parity bits
 * live in a static device array of PPH_NOALLOC_SCRATCH_BYTES per device (the
 * work is chunked to fit it; every n_in <= 32768 fits, a shape that would not
 * returns PPH_E_INVALID), and the exp table in a static device array.  The
 * first call on a device finds the arrays' addresses and fills the table (on
 * `stream`, waited for).  Every call on a device uses
 * the same arrays: calls that overlap on one GPU race, so serialise them
 * (one stream, or events between streams). */
#define PPH_NOALLOC_SCRATCH_BYTES (256ull << 20)
int pph_muladd_bilinear_noalloc(const void* in_slots, size_t slot_bytes, unsigned n_in, const uint16_t* input_logs,
                                const uint16_t* exps, void* out_slots, unsigned n_out, int accumulate, void* stream);
/* Test hook: allocations of the shared bilinear workspace above `bytes` fail
 * with PPH_E_NOMEM (0 = no limit).  The backend then runs such batches in the
 * plain kernel (the same bytes). */
int pph_debug_bilinear_workspace_limit(size_t bytes);
/* Host reference of the bilinear formula: c * x in GF(2^16) through the
 * expansion, componentwise product, recombination and reduction. */
uint16_t pph_bilinear_mul(uint16_t c, uint16_t x);

/* Per-slice hashes of the input path (hasher/hasher_base.h:83-198): for slot k,
 * MD5 (md5_out + 16k) and CRC-32 (crc_out[k]) of bytes [0, lens[k]) zero-padded
 * to slice_size (lens: device array, NULL = slice_size; slice_size % 4 == 0).
 * MD5: one lane per slice; CRC-32: up to 64 lanes per slice, combined.  Run on
 * data-form slots (16-byte aligned), before log-domain preparation. */
int pph_slice_hash(const void* slots, size_t slot_bytes, unsigned count, const uint32_t* lens, size_t slice_size,
                   uint8_t* md5_out, uint32_t* crc_out, void* stream);
/* The same hashes as IFSC entries (src/gf.cc:845-891, the input file slice
 * checksums of the PAR2 IFSC packet): slice k's MD5 then its CRC-32 (little
 * endian) at ifsc_out + 20k (device; 4-byte aligned). */
int pph_slice_hash_ifsc(const void* slots, size_t slot_bytes, unsigned count, const uint32_t* lens,
                        size_t slice_size, uint8_t* ifsc_out, void* stream);
/* MD5 of a file laid out over `count` slots, slice_size bytes each (the last
 * one partial): file_size bytes in total (hasher_base.h:178-186).  One serial
 * chain, run by one wavefront. */
int pph_file_md5(const void* slots, size_t slot_bytes, unsigned count, size_t slice_size, uint64_t file_size,
                 uint8_t* md5_out, void* stream);

/* gf16_cksum verification of n slots (copy_cksum_check, gf16/gf16_cksum_base.h:48-59):
 * valid[k] (device int array) = 1 if the stored block matches the slice data. */
int pph_verify_slots(const void* slots, size_t slot_bytes, unsigned count, size_t slice_size,
                     unsigned cksum_size, int* valid, void* stream);

/* ======================================================================
 * Batched execution: many independent jobs in one persistent launch
 * ====================================================================== */
/* One PAR2 job, device-resident.  Inputs are data-form slots on entry to every
 * run and are prepared in place (zero padding, gf16_cksum, log domain for the
 * plain form), so the caller refills them before the next run. */
typedef struct pph_job {
	void* in_slots;            /* n_in input slots of slot_bytes */
	void* out_slots;           /* n_out recovery slots of slot_bytes */
	const uint32_t* in_lens;   /* device: bytes of each input (<= slice_size); NULL = slice_size */
	const uint16_t* in_logs;   /* device: PAR2 input logs (pph_input_log of the input numbers) */
	const uint16_t* exps;      /* device: recovery exponents */
	int* valid;                /* device: n_out gf16_cksum flags, written by every run */
	size_t slice_size;
	size_t slot_bytes;         /* >= pph_slot_layout(slice_size, cksum_size) */
	unsigned n_in, n_out;      /* 1 <= n_in <= 32768 when n_out > 0 */
} pph_job;

/* Measured mul-add rates of the forms by recovery-count bucket
 * (R <= 1, 4, 16, 64, 256, 1024, 4096, more), mul-adds per second, on 64 KiB
 * slices: the plain and dense forms with 512 inputs, the bilinear form with
 * 512 (bilinear) and 4096 inputs (bilinear_wide; the planner interpolates in
 * log(inputs)).  A zero rate leaves the form out. */
typedef struct pph_tuning {
	unsigned version;
	unsigned r_buckets;
	float plain_g8[8];
	float plain_g16[8];
	float dense[8];
	float bilinear[8];
	float bilinear_wide[8];
} pph_tuning;

/* Measure the rates on the current device (about a second of GPU time). */
int pph_autotune(pph_tuning* out);

typedef struct pph_batch pph_batch;
/* Plan a batch: tuning NULL runs (once per process) and uses pph_autotune. */
int pph_batch_create(const pph_job* jobs, unsigned n_jobs, unsigned cksum_size, const pph_tuning* tuning,
                     pph_batch** out);
/* The form (PPH_FORM_PLAIN / _DENSE / _BILINEAR) chosen for a job, the plain
 * kernel's output grouping and the predicted mul-add time of the batch. */
int pph_batch_plan(const pph_batch* b, unsigned job, int* form, unsigned* group, double* predicted_s);
/* Run the batch on a stream: preparation, the fused persistent mul-add of the
 * plain jobs and the dense launches on a second stream of the batch (forked
 * from `stream` and joined back), beside the bilinear jobs' preparation and
 * launches on `stream`; then the verification of every job.  use_graph
 * captures that work into a hipGraph on first use and replays it.  A batch
 * with bilinear jobs owns a device workspace for them. */
int pph_batch_run(pph_batch* b, int use_graph, void* stream);
/* Time of the mul-add launches of the last eager (use_graph = 0) run, from HIP
 * events on its stream: from the start of the bilinear launches (which
 * prepare and verify their jobs wave by wave) to the end of every mul-add
 * launch (the plain jobs' work may start before it); waits for that run. */
int pph_batch_fused_ms(pph_batch* b, float* ms);
void pph_batch_destroy(pph_batch* b);

/* ======================================================================
 * Recovery-packet MD5 on the device: the MD5Multi of HasherOutput
 * (src/gf.cc:1070-1219) as lib/par2.js:243-262,300-327 feeds it.  Region r
 * hashes recovery packet r from its recovery set ID on (lib/par2.js:505-513):
 * set_id (16 bytes) || "PAR 2.0\0RecvSlic" || u32le(exps[r]) || the recovery
 * slice, which may arrive in pieces (chunked mode).  One lane per region.
 * ====================================================================== */
typedef struct pph_packet_md5 pph_packet_md5;
/* State for `regions` packets (1..65535) on the current device. */
int pph_packet_md5_create(unsigned regions, pph_packet_md5** out);
void pph_packet_md5_destroy(pph_packet_md5* h);
/* Start every region's message; set_id (16 bytes) and exps (regions values
 * < 65535) are host arrays, read before the call returns. */
int pph_packet_md5_begin(pph_packet_md5* h, const uint8_t* set_id, const uint16_t* exps, void* stream);
/* Append bytes [0, len) of every region's slot: slot r at slots + r *
 * slot_bytes (device, 16-byte aligned slot_bytes); len even and len + 4 <=
 * slot_bytes; slots NULL appends len zero bytes. */
int pph_packet_md5_update(pph_packet_md5* h, const void* slots, size_t slot_bytes, size_t len, void* stream);
/* Pad and write the 16-byte digests to md5_out (device, 16 x regions);
 * pph_packet_md5_begin starts the next packets. */
int pph_packet_md5_final(pph_packet_md5* h, uint8_t* md5_out, void* stream);
/* Backend: append every recovery slice's current bytes (the current slice
 * size, after the batches submitted so far; zeros when nothing was added) to
 * h, which has one region per recovery slice, and return once appended.  The
 * recovery packets' MD5 once the last chunk is in, without copying the slices
 * to the host first. */
int pph_proc_packet_md5_update(pph_proc* p, pph_packet_md5* h);

/* ======================================================================
 * Serving: jobs arrive over time and run with continuous batching
 * ====================================================================== */
/* Engine settings.  hbm_budget bytes of device memory are allocated once at
 * create and shared by the admitted jobs; a job is admitted when its memory
 * (2 x inputs + recovery slices + small arrays, in slots) and one of
 * max_compute compute streams are free. */
typedef struct pph_serve_config {
	size_t hbm_budget;
	unsigned max_compute;    /* 1..64 jobs computing at once */
	size_t md5_quantum;      /* bytes of each file's MD5 chain per hash round (multiple of 64; 0 = 8 MiB) */
	size_t staging_bytes;    /* each of the two pinned staging chunks (0 = 64 MiB) */
	size_t bilinear_workspace;  /* device bytes for the bilinear form, allocated at create beside the
	                               budget (0: plain form only).  Jobs whose predicted bilinear time is
	                               lower and whose workspace fits use it, one job's GEMMs at a time. */
	const struct pph_tuning* tuning;  /* rates for that choice; NULL runs pph_autotune at create */
	unsigned md5_host_threads;  /* > 0: this many host threads compute every file MD5, beside the device work
	                               (one core runs an MD5 chain several times faster than one wavefront):
	                               host_data files from submit on, device-generated files once their
	                               inputs are in HBM, read back through two pinned chunks per thread of
	                               min(staging_bytes, 4 MiB); 0: every file MD5 on the device */
	unsigned upload_threads;    /* threads copying host-sourced inputs to the device, each with its own two
	                               pinned staging chunks (0 = 1, at most 16) */
	unsigned md5_order;         /* which queued file a free host MD5 thread takes: 0 = highest response
	                               ratio, (wait + expected hash time) / expected hash time, so short files
	                               do not queue behind multi-second chains while long ones still age to
	                               the front; 1 = submit order */
} pph_serve_config;

/* One file.  Its slices (slice_size bytes, the last one partial) are PAR2
 * inputs first_input, first_input + 1, ...; the job computes, for every slice,
 * MD5 and CRC-32 of the slice zero-padded to slice_size (the IFSC entries,
 * hasher/hasher_base.h:154-176), the MD5 of the file (hasher_base.h:178-186),
 * and n_out recovery slices with their gf16_cksum flags (PAR2Proc over the
 * file's slices, gf16/controller.h:248-326), and on request each recovery
 * packet's MD5.  Output pointers are host memory, may be NULL, and are
 * written before the job's result is posted. */
typedef struct pph_serve_job {
	uint64_t file_size;         /* >= 1 */
	size_t slice_size;          /* multiple of 4; ceil(file_size / slice_size) + first_input <= 32768 */
	unsigned n_out;             /* recovery slices (0: hashes only) */
	unsigned first_input;
	const uint16_t* exps;       /* n_out exponents (< 65535), copied at submit; NULL = 0 .. n_out-1 */
	const void* host_data;      /* file bytes, read until the job completes; NULL = generated on the
	                               device: slice k = pph_fill_random(seed, job_id, first slice k) */
	uint64_t seed;
	uint32_t job_id;
	void* recovery_out;         /* n_out x slice_size bytes */
	int* valid;                 /* n_out flags */
	uint8_t* slice_md5;         /* 16 bytes per slice */
	uint32_t* slice_crc;        /* one per slice */
	uint8_t* file_md5;          /* 16 bytes */
	uint8_t* ifsc;              /* 20 bytes per slice: MD5 then CRC-32, the IFSC entries (NULL: none) */
	const uint8_t* packet_set_id;  /* 16 bytes, copied at submit: the recovery set ID; with it, each
	                                  recovery packet's MD5 (lib/par2.js:505-513) goes to packet_md5 */
	uint8_t* packet_md5;        /* 16 bytes per recovery slice */
} pph_serve_job;

typedef struct pph_serve_result {
	uint64_t ticket;
	int status;                 /* PPH_OK or the error that ended the job */
	double latency_s;           /* submit -> every output written */
	double queue_s;             /* submit -> admitted */
	double compute_s;           /* admitted -> recovery, flags and slice hashes written */
	double run_s;               /* admitted -> done (includes the file MD5 chain) */
	double mul_adds;            /* n_in x n_out x slice_size / 2 */
	double bytes;               /* file_size */
} pph_serve_result;

typedef struct pph_serve_stats {
	uint64_t submitted, completed, md5_rounds;
	uint64_t bilinear_jobs;     /* completed jobs that used the bilinear form */
	uint64_t host_md5_jobs;     /* completed jobs whose file MD5 ran on the host */
	unsigned peak_jobs;         /* most jobs admitted at once */
	uint64_t peak_bytes;        /* most arena bytes in use at once */
	uint64_t arena_bytes;
	double mul_adds, bytes;     /* of the jobs completed without error */
} pph_serve_stats;

typedef struct pph_serve pph_serve;
int pph_serve_create(int device, const pph_serve_config* cfg, pph_serve** out);
/* Queue a job (thread-safe).  PPH_E_NOMEM: the job can never fit the budget. */
int pph_serve_submit(pph_serve* s, const pph_serve_job* job, uint64_t* ticket);
/* Up to max finished jobs' results, waiting up to wait_ms for the first. */
int pph_serve_poll(pph_serve* s, pph_serve_result* out, unsigned max, unsigned* n, int wait_ms);
/* Wait up to wait_ms for every submitted job to finish; returns how many have not. */
int pph_serve_drain(pph_serve* s, int wait_ms);
int pph_serve_get_stats(pph_serve* s, pph_serve_stats* out);
/* Finishes every submitted job, then frees the engine. */
void pph_serve_destroy(pph_serve* s);

#ifdef __cplusplus
}
#endif
#endif

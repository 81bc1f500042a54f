// Batched execution: many independent PAR2 jobs (uneven input counts, slice
// sizes and recovery counts): the plain jobs in one persistent fused mul-add
// launch, the bilinear jobs in cross-job waves, with the preparation and
// checksum verification around them (one launch each) captured once in a
// hipGraph and replayed.  The form of each job (plain, dense bit-matrix or
// bilinear) and the plain kernel's output grouping are chosen per batch from
// rates the autotuner measures on the device.
#include "pph_internal.h"
#include "pph_plain_tile.h"
#include "../../include/parpar_hip.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

namespace pph {

struct BatchJob {
	const uint16_t* in;
	uint16_t* out;
	const uint16_t* in_logs;
	const uint16_t* exps;
	uint64_t slot_words;
	uint32_t n_in, n_out;
};

struct BatchItem {
	uint32_t job;
	uint32_t grp;
	uint32_t col;
	uint32_t pad;
};

// Persistent: one workgroup per CU loads the exp table once, then takes work
// items (job, column tile, output group) from an atomic counter until they run
// out.  Every wave leaves the loop on the same (LDS-broadcast) item index.
template <int G>
__global__ __launch_bounds__(kPlainThreads) void k_muladd_batch(const BatchJob* __restrict__ jobs,
                                                                const BatchItem* __restrict__ items, unsigned n_items,
                                                                unsigned* __restrict__ counter,
                                                                const uint16_t* __restrict__ exp_tab) {
	__shared__ uint16_t s_exp[65536];
	__shared__ uint16_t s_lin[kPlainThreads];
	__shared__ unsigned s_item;
	load_exp_table(s_exp, exp_tab);
	for(;;) {
		if(threadIdx.x == 0) s_item = atomicAdd(counter, 1u);
		__syncthreads();
		const unsigned k = s_item;
		__syncthreads();  // every thread has read s_item before it is overwritten
		if(k >= n_items) break;
		const BatchItem it = items[k];
		const BatchJob jb = jobs[it.job];
		plain_tile<G, false>(s_exp, s_lin, jb.in, jb.slot_words, jb.n_in, jb.in_logs, jb.exps, nullptr, 0, jb.out, jb.n_out,
		                     0, it.col, it.grp);
	}
}

__global__ void k_set_u32(unsigned* p, unsigned v) {
	if(threadIdx.x == 0) *p = v;
}

// ------------------------------------------------------------------ autotuner

constexpr unsigned kTuneBuckets = 8;
static const unsigned kTuneR[kTuneBuckets] = {1, 4, 16, 64, 256, 1024, 4096, 8192};

unsigned tune_bucket(unsigned r) {
	for(unsigned b = 0; b < kTuneBuckets; b++)
		if(r <= kTuneR[b]) return b;
	return kTuneBuckets - 1;
}

// Share of the bilinear form's padded work that is real: its GEMM pads the
// inputs to a multiple of 256 and the outputs to its row tile (blm_row_tile;
// in a batch, to the weighed rows of the job's mixed row tiles).
double bilinear_fill(unsigned n_in, unsigned n_out, bool batched) {
	const double rows = batched ? blm_row_split(n_out).cost : (double)round_up(n_out, blm_row_tile(n_out));
	return (double)n_in / (double)round_up(n_in, 256) * ((double)n_out / rows);
}

double bilinear_rate(const pph_tuning& tu, unsigned n_in, unsigned n_out, bool batched) {
	const unsigned bk = tune_bucket(n_out);
	if(tu.version < 3 || !(tu.bilinear[bk] > 0) || !(tu.bilinear_wide[bk] > 0) || !n_in) return 0;
	// log-linear in the input count between the two measured widths (512, 4096)
	const double f = std::min(1.0, std::max(0.0, (std::log2((double)n_in) - 9.0) / 3.0));
	const double v = std::exp((1 - f) * std::log((double)tu.bilinear[bk]) + f * std::log((double)tu.bilinear_wide[bk]));
	// scaled by the job's padding against that of the measured shapes
	return v * bilinear_fill(n_in, n_out, batched) / bilinear_fill(512, kTuneR[bk], batched);
}

double plain_rate(const pph_tuning& tu, unsigned n_out) {
	const unsigned bk = tune_bucket(n_out);
	return std::max(1.0f, std::max(tu.plain_g8[bk], tu.plain_g16[bk]));
}

int autotune(pph_tuning* out) {
	DeviceTables* t;
	int rc = device_tables(&t);
	if(rc) return rc;
	const unsigned N = 512, N_wide = 4096;
	const size_t S = 65536;
	const unsigned ck = 64;
	const size_t slot = slot_size(S, ck);
	const unsigned R_max = kTuneR[kTuneBuckets - 1];
	uint8_t *in_log = nullptr, *in_dat = nullptr, *outb = nullptr;
	uint16_t *logs = nullptr, *exps = nullptr;
	hipStream_t s = nullptr;
	hipEvent_t e0 = nullptr, e1 = nullptr;
	auto cleanup = [&]() {
		if(in_log) (void)hipFree(in_log);
		if(in_dat) (void)hipFree(in_dat);
		if(outb) (void)hipFree(outb);
		if(logs) (void)hipFree(logs);
		if(exps) (void)hipFree(exps);
		if(e0) (void)hipEventDestroy(e0);
		if(e1) (void)hipEventDestroy(e1);
		if(s) (void)hipStreamDestroy(s);
	};
	if(hipMalloc(&in_log, N * slot) != hipSuccess || hipMalloc(&in_dat, N_wide * slot) != hipSuccess ||
	   hipMalloc(&outb, (size_t)1 << 30) != hipSuccess || hipMalloc(&logs, N_wide * 2) != hipSuccess ||
	   hipMalloc(&exps, R_max * 2) != hipSuccess) {
		cleanup();
		return PPH_E_NOMEM;
	}
	std::vector<uint16_t> h(std::max(N_wide, R_max));
	hipError_t e = hipSuccess;
	for(unsigned i = 0; i < N_wide; i++) h[i] = host_input_log((uint16_t)i);
	if(e == hipSuccess) e = hipMemcpy(logs, h.data(), N_wide * 2, hipMemcpyHostToDevice);
	for(unsigned r = 0; r < R_max; r++) h[r] = (uint16_t)r;
	if(e == hipSuccess) e = hipMemcpy(exps, h.data(), R_max * 2, hipMemcpyHostToDevice);
	if(e == hipSuccess) e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
	if(e == hipSuccess) e = hipEventCreate(&e0);
	if(e == hipSuccess) e = hipEventCreate(&e1);
	if(e == hipSuccess) e = launch_fill_random(in_dat, slot, N, S, 12345, 0, 0, s);
	if(e == hipSuccess) e = launch_prep(in_log, in_dat, slot, N, nullptr, nullptr, S, ck_offset(S), ck, true, *t, s);
	if(e == hipSuccess) e = launch_fill_random(in_dat, slot, N_wide, S, 12345, 0, 0, s);
	if(e == hipSuccess) e = launch_prep(in_dat, nullptr, slot, N_wide, nullptr, nullptr, S, ck_offset(S), ck, false, *t, s);
	// The plain form runs as the persistent batched kernel: measure it that way,
	// on J jobs that share the input slots, enough work items to fill the GPU.
	BatchJob* d_jobs = nullptr;
	BatchItem* d_items = nullptr;
	unsigned* d_counter = nullptr;
	const size_t out_cap = (size_t)1 << 30;
	auto timed = [&](int form, unsigned R, float* ms, double* ma, unsigned n_in) -> hipError_t {
		hipError_t r = hipSuccess;
		const unsigned G = form == 0 ? 8 : 16;
		const unsigned tile_words = kPlainThreads * kPlainWordsPerLane;
		const unsigned cols = (unsigned)((slot / 2 + tile_words - 1) / tile_words);
		const unsigned per_job = cols * ((R + G - 1) / G);
		unsigned J = (2048 + per_job - 1) / per_job;
		const unsigned jmax = (unsigned)std::max<size_t>(1, out_cap / ((size_t)R * slot));
		if(form >= 2) J = 1;
		J = std::min(J, jmax);
		*ma = (double)J * n_in * R * (S / 2);
		if(form < 2) {  // plain: the persistent batched kernel
			std::vector<BatchJob> hj(J);
			std::vector<BatchItem> hi;
			for(unsigned j = 0; j < J; j++) {
				hj[j] = BatchJob{(const uint16_t*)in_log, (uint16_t*)(outb + (size_t)j * R * slot), logs, exps, slot / 2, N, R};
				for(unsigned c = 0; c < cols; c++)
					for(unsigned gp = 0; gp < (R + G - 1) / G; gp++) hi.push_back(BatchItem{j, gp, c, 0});
			}
			if(d_jobs) (void)hipFree(d_jobs);
			if(d_items) (void)hipFree(d_items);
			d_jobs = nullptr;
			d_items = nullptr;
			r = hipMalloc(&d_jobs, J * sizeof(BatchJob));
			if(r == hipSuccess) r = hipMalloc(&d_items, hi.size() * sizeof(BatchItem));
			if(r == hipSuccess) r = hipMemcpy(d_jobs, hj.data(), J * sizeof(BatchJob), hipMemcpyHostToDevice);
			if(r == hipSuccess) r = hipMemcpy(d_items, hi.data(), hi.size() * sizeof(BatchItem), hipMemcpyHostToDevice);
			for(int rep = 0; rep < 2 && r == hipSuccess; rep++) {
				hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(64), 0, s, d_counter, 0u);
				r = hipEventRecord(e0, s);
				if(r != hipSuccess) break;
				dim3 grid(t->cus > 0 ? t->cus : 256), block(kPlainThreads);
				if(G == 8)
					hipLaunchKernelGGL(k_muladd_batch<8>, grid, block, 0, s, d_jobs, d_items, (unsigned)hi.size(), d_counter,
					                   t->exp);
				else
					hipLaunchKernelGGL(k_muladd_batch<16>, grid, block, 0, s, d_jobs, d_items, (unsigned)hi.size(),
					                   d_counter, t->exp);
				r = hipGetLastError();
				if(r == hipSuccess) r = hipEventRecord(e1, s);
				if(r == hipSuccess) r = hipEventSynchronize(e1);
				if(r == hipSuccess) r = hipEventElapsedTime(ms, e0, e1);
			}
			return r;
		}
		for(int rep = 0; rep < 2 && r == hipSuccess; rep++) {
			r = hipEventRecord(e0, s);
			if(r != hipSuccess) break;
			if(form == 3) {
				const int st = pph_muladd_bilinear(in_dat, slot, n_in, logs, exps, outb, R, 0, s);
				if(st) {
					*ma = 0;  // form unavailable: rate 0
					return hipSuccess;
				}
			} else {
				r = launch_muladd_dense((const uint16_t*)in_dat, slot / 2, N, logs, exps, (uint16_t*)outb, R, false, *t, s);
			}
			if(r == hipSuccess) r = hipEventRecord(e1, s);
			if(r == hipSuccess) r = hipEventSynchronize(e1);
			if(r == hipSuccess) r = hipEventElapsedTime(ms, e0, e1);
		}
		return r;
	};
	if(e == hipSuccess) e = hipMalloc(&d_counter, 256);
	pph_tuning tu;
	memset(&tu, 0, sizeof(tu));
	tu.version = 3;
	tu.r_buckets = kTuneBuckets;
	for(unsigned b = 0; b < kTuneBuckets && e == hipSuccess; b++) {
		const unsigned R = kTuneR[b];
		double ma = 0;
		float ms = 0;
		e = timed(0, R, &ms, &ma, N);
		tu.plain_g8[b] = (float)(ma / (ms * 1e-3));
		if(e == hipSuccess) e = timed(1, R, &ms, &ma, N);
		tu.plain_g16[b] = (float)(ma / (ms * 1e-3));
		if(e == hipSuccess) e = timed(2, R, &ms, &ma, N);
		tu.dense[b] = (float)(ma / (ms * 1e-3));
		if(e == hipSuccess) e = timed(3, R, &ms, &ma, N);
		tu.bilinear[b] = ma > 0 ? (float)(ma / (ms * 1e-3)) : 0.0f;
		if(e == hipSuccess) e = timed(3, R, &ms, &ma, N_wide);
		tu.bilinear_wide[b] = ma > 0 ? (float)(ma / (ms * 1e-3)) : 0.0f;
	}
	if(d_jobs) (void)hipFree(d_jobs);
	if(d_items) (void)hipFree(d_items);
	if(d_counter) (void)hipFree(d_counter);
	cleanup();
	(void)pph_bilinear_release();
	if(e != hipSuccess) return set_error(e);
	*out = tu;
	return PPH_OK;
}

}  // namespace pph

using namespace pph;

static std::mutex g_tune_mu;
static bool g_tune_valid = false;
static pph_tuning g_tune;

extern "C" int pph_autotune(pph_tuning* out) {
	if(!out) return PPH_E_INVALID;
	int n = 0;
	if(hipGetDeviceCount(&n) != hipSuccess || n <= 0) return PPH_E_UNSUPPORTED;
	std::lock_guard<std::mutex> lk(g_tune_mu);
	int rc = autotune(out);
	if(rc == PPH_OK) {
		g_tune = *out;
		g_tune_valid = true;
	}
	return rc;
}

// ====================================================================== batch

struct pph_batch {
	int device = 0;
	unsigned ck = 64;
	std::vector<pph_job> jobs;
	std::vector<int> form;         // PPH_FORM_PLAIN / _DENSE / _BILINEAR per job
	void* bl_ws = nullptr;         // workspace of the bilinear jobs
	size_t bl_ws_bytes = 0;
	BlmBatch* bl_plan = nullptr;   // the bilinear jobs run together (cross-job launches)
	std::vector<char> bl_alone;    // bilinear jobs outside the plan (run one by one)
	unsigned group = 16;           // plain kernel output grouping
	DeviceTables* tab = nullptr;
	BatchJob* d_jobs = nullptr;
	BatchItem* d_items = nullptr;
	unsigned n_items = 0;
	unsigned* d_counter = nullptr;
	hipStream_t cap = nullptr;
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	double predicted_s = 0;
	hipEvent_t ev_a = nullptr, ev_b = nullptr;  // around the mul-add launches (eager runs)
	bool timing = false;
	// second stream of the plain and dense jobs, forked from and joined back
	// to the caller's stream
	hipStream_t aux = nullptr;
	hipEvent_t ev_fork = nullptr, ev_join = nullptr;
	// one-launch preparation and verification tables (device): the bilinear
	// jobs' preparation, the dense jobs', then every job's verification
	CkJob* d_ck = nullptr;
	unsigned ck_n[3] = {0, 0, 0}, ck_max[3] = {0, 0, 0};  // entries; slots in all, or the largest count
	const CkJob* ck_table(unsigned k) const { return d_ck + (k > 0 ? ck_n[0] : 0) + (k > 1 ? ck_n[1] : 0); }
};

// Bilinear jobs' mul-adds: those of the plan in its cross-job launches, the
// others (too large for its workspace share) one by one.
static hipError_t enqueue_bilinear(pph_batch* b, hipStream_t s) {
	if(b->bl_plan) {
		const int st = blm_batch_enqueue(b->bl_plan, *b->tab, s);
		if(st) return hipErrorLaunchFailure;
	}
	for(size_t j = 0; j < b->jobs.size(); j++) {
		const pph_job& jb = b->jobs[j];
		if(b->form[j] != PPH_FORM_BILINEAR || !jb.n_out || !b->bl_alone[j]) continue;
		const int st = bilinear_enqueue(jb.in_slots, jb.slot_bytes, jb.n_in, jb.in_logs, jb.exps, jb.out_slots, jb.n_out,
		                                false, b->bl_ws, b->bl_ws_bytes, *b->tab, s);
		if(st) return hipErrorLaunchFailure;
	}
	return hipSuccess;
}

// The batch's device work, on two streams: the plain and dense jobs (their
// preparation, the dense launches, the persistent plain launch) on `aux`,
// beside the bilinear jobs' preparation and launches on the caller's stream,
// so that neither waits for the other.  `aux` joins back before the
// verification of every job (one launch).
static hipError_t enqueue_batch(pph_batch* b, hipStream_t s) {
	hipError_t e = hipEventRecord(b->ev_fork, s);
	if(e == hipSuccess) e = hipStreamWaitEvent(b->aux, b->ev_fork, 0);
	if(e != hipSuccess) return e;
	const hipStream_t a = b->aux;
	if(b->n_items) {
		hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(64), 0, a, b->d_counter, 0u);
		e = hipGetLastError();
	}
	// the dense jobs first: short launches that would hold CUs for tens of ms
	// beside the bilinear GEMMs (a dense workgroup and a GEMM workgroup do not
	// fit on one CU); then the plain jobs, whose launch runs as CUs free up
	if(e == hipSuccess) e = launch_prep_data_jobs(b->ck_table(1), b->ck_n[1], b->ck_max[1], b->ck, 8, *b->tab, a);
	for(size_t j = 0; j < b->jobs.size() && e == hipSuccess; j++) {
		const pph_job& jb = b->jobs[j];
		if(b->form[j] != PPH_FORM_DENSE || !jb.n_out) continue;
		e = launch_muladd_dense((const uint16_t*)jb.in_slots, jb.slot_bytes / 2, jb.n_in, jb.in_logs, jb.exps,
		                        (uint16_t*)jb.out_slots, jb.n_out, false, *b->tab, a);
	}
	for(size_t j = 0; j < b->jobs.size() && e == hipSuccess; j++) {
		const pph_job& jb = b->jobs[j];
		if(b->form[j] != PPH_FORM_PLAIN || !jb.n_in || !jb.n_out) continue;
		e = launch_prep((uint8_t*)jb.in_slots, nullptr, jb.slot_bytes, jb.n_in, jb.in_lens, nullptr, jb.slice_size,
		                ck_offset(jb.slice_size), b->ck, true, *b->tab, a);
	}
	if(e == hipSuccess && b->n_items) {
		dim3 grid(b->tab->cus > 0 ? b->tab->cus : 256), block(kPlainThreads);
		if(b->group == 8)
			hipLaunchKernelGGL(k_muladd_batch<8>, grid, block, 0, a, b->d_jobs, b->d_items, b->n_items, b->d_counter,
			                   b->tab->exp);
		else
			hipLaunchKernelGGL(k_muladd_batch<16>, grid, block, 0, a, b->d_jobs, b->d_items, b->n_items, b->d_counter,
			                   b->tab->exp);
		e = hipGetLastError();
	}
	// the join is recorded even after an error, so that a capture on `s` can end
	const hipError_t ej = hipEventRecord(b->ev_join, a);
	// Bilinear jobs outside the plan (the plan prepares its own jobs wave by
	// wave): their preparation on the whole device (eight workgroups per CU;
	// two per CU leave room for the plain launches beside them but move the
	// bytes at half the rate: 500 against 489 ms per replay on the 64-job
	// workload, when all its bilinear jobs were prepared here).
	hipError_t eb = launch_prep_data_jobs(b->ck_table(0), b->ck_n[0], b->ck_max[0], b->ck, 8, *b->tab, s);
	if(eb == hipSuccess && b->timing) eb = hipEventRecord(b->ev_a, s);
	if(eb == hipSuccess) eb = enqueue_bilinear(b, s);
	const hipError_t ew = ej == hipSuccess ? hipStreamWaitEvent(s, b->ev_join, 0) : ej;
	if(e == hipSuccess) e = eb;
	if(e == hipSuccess) e = ew;
	if(e == hipSuccess && b->timing) e = hipEventRecord(b->ev_b, s);
	if(e == hipSuccess) e = launch_verify_jobs(b->ck_table(2), b->ck_n[2], b->ck_max[2], b->ck, *b->tab, s);
	return e;
}

extern "C" void pph_batch_destroy(pph_batch* b) {
	if(!b) return;
	(void)hipSetDevice(b->device);
	if(b->cap) (void)hipStreamSynchronize(b->cap);
	if(b->exec) (void)hipGraphExecDestroy(b->exec);
	if(b->graph) (void)hipGraphDestroy(b->graph);
	if(b->cap) (void)hipStreamDestroy(b->cap);
	if(b->d_jobs) (void)hipFree(b->d_jobs);
	if(b->d_items) (void)hipFree(b->d_items);
	if(b->d_counter) (void)hipFree(b->d_counter);
	if(b->aux) (void)hipStreamSynchronize(b->aux);
	if(b->ev_a) (void)hipEventDestroy(b->ev_a);
	if(b->ev_b) (void)hipEventDestroy(b->ev_b);
	if(b->ev_fork) (void)hipEventDestroy(b->ev_fork);
	if(b->ev_join) (void)hipEventDestroy(b->ev_join);
	if(b->aux) (void)hipStreamDestroy(b->aux);
	if(b->d_ck) (void)hipFree(b->d_ck);
	blm_batch_destroy(b->bl_plan);
	if(b->bl_ws) (void)hipFree(b->bl_ws);
	delete b;
}

extern "C" int pph_batch_create(const pph_job* jobs, unsigned n_jobs, unsigned cksum_size, const pph_tuning* tuning,
                                pph_batch** out) {
	if(!out || (!jobs && n_jobs) || !valid_ck(cksum_size)) return PPH_E_INVALID;
	*out = nullptr;
	for(unsigned j = 0; j < n_jobs; j++) {
		const pph_job& jb = jobs[j];
		if(jb.n_in > 32768 || jb.n_out > 65535) return PPH_E_INVALID;
		if(jb.slot_bytes % kSlotAlign || jb.slot_bytes < slot_size(jb.slice_size, cksum_size)) return PPH_E_INVALID;
		if(jb.n_out && (!jb.out_slots || !jb.valid)) return PPH_E_INVALID;
		if(jb.n_out && (!jb.n_in || !jb.in_slots || !jb.in_logs || !jb.exps)) return PPH_E_INVALID;
	}
	int dev = 0;
	if(hipGetDevice(&dev) != hipSuccess) return PPH_E_UNSUPPORTED;
	pph_tuning tu;
	if(tuning) {
		tu = *tuning;
	} else {
		std::lock_guard<std::mutex> lk(g_tune_mu);
		if(!g_tune_valid) {
			int rc = autotune(&g_tune);
			if(rc) return rc;
			g_tune_valid = true;
		}
		tu = g_tune;
	}
	auto* b = new(std::nothrow) pph_batch;
	if(!b) return PPH_E_NOMEM;
	b->device = dev;
	b->ck = cksum_size;
	b->jobs.assign(jobs, jobs + n_jobs);
	int rc = device_tables(&b->tab);
	if(rc) {
		pph_batch_destroy(b);
		return rc;
	}

	// choose the plain grouping for the batch, then the form per job, from the
	// measured rates (mul-adds per second by recovery-count bucket)
	auto ma_of = [](const pph_job& jb) { return (double)jb.n_in * jb.n_out * (double)(jb.slot_bytes / 2); };
	// The plain kernel computes whole groups of G rows (the rows past the
	// recovery count run with exponent 0 and are not stored), so a job costs
	// its rows rounded up to G at the rate measured for that many rows.
	auto plain_time = [&](const pph_job& jb, unsigned G) {
		const unsigned rows = (jb.n_out + G - 1) / G * G;
		const unsigned bk = tune_bucket(rows);
		return ma_of(jb) * ((double)rows / jb.n_out) / std::max(1.0f, G == 8 ? tu.plain_g8[bk] : tu.plain_g16[bk]);
	};
	double t8 = 0, t16 = 0;
	for(const auto& jb : b->jobs) {
		if(!jb.n_in || !jb.n_out) continue;
		t8 += plain_time(jb, 8);
		t16 += plain_time(jb, 16);
	}
	b->group = t8 < t16 ? 8 : 16;
	std::vector<double> pred(n_jobs, 0.0);
	b->form.assign(n_jobs, PPH_FORM_PLAIN);
	for(unsigned j = 0; j < n_jobs; j++) {
		const pph_job& jb = b->jobs[j];
		if(!jb.n_in || !jb.n_out) continue;
		const unsigned bk = tune_bucket(jb.n_out);
		const double ma = ma_of(jb);
		const double tp = plain_time(jb, b->group);
		const double td = tu.dense[bk] > 0 ? ma / tu.dense[bk] : 1e30;
		const double rb = bilinear_rate(tu, jb.n_in, jb.n_out, true);
		const double tb = rb > 0 ? ma / rb : 1e30;
		double best = tp;
		if(td < best) {
			best = td;
			b->form[j] = PPH_FORM_DENSE;
		}
		if(tb < best && bilinear_supported(jb.slot_bytes, jb.n_in, jb.n_out) == PPH_OK) {
			best = tb;
			b->form[j] = PPH_FORM_BILINEAR;
		}
		pred[j] = best;
		b->predicted_s += pred[j];
	}
	// The bilinear jobs' workspace: all of their pieces at once where that fits
	// a budget (96 GiB, at most the free memory less 8 GiB), else waves of
	// them; jobs whose planes would take more than a quarter of it run alone.
	// (The 64-job workload: 451-454 ms per replay with 96 or 160 GiB, 461 with
	// 48 and 471 with 24: fewer waves.)
	size_t budget = (size_t)96 << 30;
	{
		size_t fr = 0, tot = 0;
		if(hipMemGetInfo(&fr, &tot) == hipSuccess && fr > ((size_t)8 << 30))
			budget = std::min(budget, fr - ((size_t)8 << 30));
		if(const char* v = getenv("PPH_BATCH_BILINEAR_WORKSPACE")) budget = (size_t)strtoull(v, nullptr, 10);
	}
	std::vector<BlmBatchJob> bl;
	b->bl_alone.assign(n_jobs, 0);
	size_t alone_ws = 0;
	for(unsigned j = 0; j < n_jobs; j++) {
		if(b->form[j] != PPH_FORM_BILINEAR) continue;
		const pph_job& jb = b->jobs[j];
		const BlmBatchJob x{jb.in_slots, jb.slot_bytes, jb.n_in,       jb.in_logs, jb.exps,
		                    jb.out_slots, jb.n_out,     jb.in_lens, jb.slice_size, jb.valid};
		if(blm_batch_eligible(x, budget)) {
			bl.push_back(x);
		} else {
			b->bl_alone[j] = 1;
			alone_ws = std::max(alone_ws, bilinear_workspace_bytes(jb.slot_bytes, jb.n_in, jb.n_out));
		}
	}
	b->bl_ws_bytes = std::max(alone_ws, blm_batch_workspace(bl.data(), (unsigned)bl.size(), budget));
	if(b->bl_ws_bytes && hipMalloc(&b->bl_ws, b->bl_ws_bytes) != hipSuccess) {
		(void)hipGetLastError();
		b->bl_ws = nullptr;
		b->bl_ws_bytes = 0;
		pph_batch_destroy(b);
		return PPH_E_NOMEM;
	}
	if(!bl.empty()) {
		rc = blm_batch_create(bl.data(), (unsigned)bl.size(), b->bl_ws, b->bl_ws_bytes, b->ck, &b->bl_plan);
		if(rc) {
			pph_batch_destroy(b);
			return rc;
		}
	}

	// work items of the plain jobs, longest job first (largest first keeps the
	// tail of the persistent launch short); within a job, output groups of one
	// column tile are adjacent so their shared input reads meet in L2
	std::vector<unsigned> order;
	for(unsigned j = 0; j < n_jobs; j++)
		if(b->form[j] == PPH_FORM_PLAIN && b->jobs[j].n_in && b->jobs[j].n_out) order.push_back(j);
	std::stable_sort(order.begin(), order.end(), [&](unsigned a, unsigned c) { return pred[a] > pred[c]; });
	std::vector<BatchItem> items;
	std::vector<BatchJob> hj(n_jobs);
	const unsigned tile_words = kPlainThreads * kPlainWordsPerLane;
	for(unsigned j = 0; j < n_jobs; j++) {
		const pph_job& jb = b->jobs[j];
		hj[j] = BatchJob{(const uint16_t*)jb.in_slots, (uint16_t*)jb.out_slots, jb.in_logs, jb.exps,
		                 jb.slot_bytes / 2, jb.n_in, jb.n_out};
	}
	for(unsigned j : order) {
		const pph_job& jb = b->jobs[j];
		const unsigned cols = (unsigned)((jb.slot_bytes / 2 + tile_words - 1) / tile_words);
		const unsigned grps = (jb.n_out + b->group - 1) / b->group;
		for(unsigned c = 0; c < cols; c++)
			for(unsigned gp = 0; gp < grps; gp++) items.push_back(BatchItem{j, gp, c, 0});
	}
	b->n_items = (unsigned)items.size();
	// preparation tables (largest slots first: the longest workgroups start
	// first) and the verification table
	std::vector<CkJob> ck_tab[3];
	for(unsigned j = 0; j < n_jobs; j++) {
		const pph_job& jb = b->jobs[j];
		if(!jb.n_in || !jb.n_out) continue;
		const CkJob prep{(uint8_t*)jb.in_slots, jb.in_lens, nullptr, jb.slot_bytes, jb.slice_size,
		                 ck_offset(jb.slice_size), jb.n_in, 0u};
		const CkJob ver{(uint8_t*)jb.out_slots, nullptr, jb.valid, jb.slot_bytes, jb.slice_size,
		                ck_offset(jb.slice_size), jb.n_out, 0u};
		if(!ck_job_valid(prep, b->ck) || !ck_job_valid(ver, b->ck)) {
			pph_batch_destroy(b);
			return PPH_E_INVALID;
		}
		// the plan prepares and verifies its own jobs wave by wave
		const bool planned = b->form[j] == PPH_FORM_BILINEAR && !b->bl_alone[j] && b->bl_plan;
		if(b->form[j] == PPH_FORM_BILINEAR && !planned) ck_tab[0].push_back(prep);
		if(b->form[j] == PPH_FORM_DENSE) ck_tab[1].push_back(prep);
		if(!planned) ck_tab[2].push_back(ver);
	}
	// ck_max: slots in all (preparation), the largest count (verification)
	std::vector<CkJob> ck_all;
	for(unsigned k = 0; k < 3; k++) {
		if(k < 2)
			std::stable_sort(ck_tab[k].begin(), ck_tab[k].end(),
			                 [](const CkJob& x, const CkJob& y) { return x.slot_bytes > y.slot_bytes; });
		b->ck_n[k] = (unsigned)ck_tab[k].size();
		for(CkJob& c : ck_tab[k]) {
			c.first = k < 2 ? b->ck_max[k] : 0u;
			b->ck_max[k] = k < 2 ? b->ck_max[k] + c.count : std::max(b->ck_max[k], c.count);
		}
		ck_all.insert(ck_all.end(), ck_tab[k].begin(), ck_tab[k].end());
	}
	hipError_t e = hipSuccess;
	if(!ck_all.empty()) e = hipMalloc(&b->d_ck, ck_all.size() * sizeof(CkJob));
	if(e == hipSuccess && !ck_all.empty())
		e = hipMemcpy(b->d_ck, ck_all.data(), ck_all.size() * sizeof(CkJob), hipMemcpyHostToDevice);
	if(e == hipSuccess && n_jobs) e = hipMalloc(&b->d_jobs, n_jobs * sizeof(BatchJob));
	if(e == hipSuccess && b->n_items) e = hipMalloc(&b->d_items, items.size() * sizeof(BatchItem));
	if(e == hipSuccess) e = hipMalloc(&b->d_counter, 256);
	if(e == hipSuccess && n_jobs) e = hipMemcpy(b->d_jobs, hj.data(), n_jobs * sizeof(BatchJob), hipMemcpyHostToDevice);
	if(e == hipSuccess && b->n_items)
		e = hipMemcpy(b->d_items, items.data(), items.size() * sizeof(BatchItem), hipMemcpyHostToDevice);
	if(e == hipSuccess) e = hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking);
	if(e == hipSuccess) e = hipEventCreate(&b->ev_a);
	if(e == hipSuccess) e = hipEventCreate(&b->ev_b);
	if(e == hipSuccess) e = hipStreamCreateWithFlags(&b->aux, hipStreamNonBlocking);
	if(e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming);
	if(e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming);
	if(e != hipSuccess) {
		pph_batch_destroy(b);
		return e == hipErrorOutOfMemory ? PPH_E_NOMEM : set_error(e);
	}
	*out = b;
	return PPH_OK;
}

extern "C" int pph_batch_plan(const pph_batch* b, unsigned job, int* form, unsigned* group, double* predicted_s) {
	if(!b || job >= b->jobs.size()) return PPH_E_INVALID;
	if(form) *form = b->form[job];
	if(group) *group = b->group;
	if(predicted_s) *predicted_s = b->predicted_s;
	return PPH_OK;
}

extern "C" int pph_batch_run(pph_batch* b, int use_graph, void* stream) {
	if(!b) return PPH_E_INVALID;
	hipStream_t s = (hipStream_t)stream;
	if(!use_graph) {
		b->timing = true;
		hipError_t e = enqueue_batch(b, s);
		b->timing = false;
		if(e != hipSuccess) return set_error(e);
		return PPH_OK;
	}
	if(!b->exec) {
		hipError_t e = hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal);
		if(e != hipSuccess) return set_error(e);
		hipError_t ee = enqueue_batch(b, b->cap);
		e = hipStreamEndCapture(b->cap, &b->graph);
		if(ee != hipSuccess) return set_error(ee);
		if(e != hipSuccess) return set_error(e);
		e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
		if(e != hipSuccess) return set_error(e);
	}
	hipError_t e = hipGraphLaunch(b->exec, s);
	if(e != hipSuccess) return set_error(e);
	return PPH_OK;
}

extern "C" int pph_batch_fused_ms(pph_batch* b, float* ms) {
	if(!b || !ms) return PPH_E_INVALID;
	hipError_t e = hipEventSynchronize(b->ev_b);
	if(e == hipSuccess) e = hipEventElapsedTime(ms, b->ev_a, b->ev_b);
	if(e != hipSuccess) return set_error(e);
	return PPH_OK;
}

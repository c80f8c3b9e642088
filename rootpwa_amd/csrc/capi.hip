// capi.hip -- likelihood contexts behind the C ABI of rootpwa_amd.h.
//
// A context owns the device copy of the decay amplitudes of one likelihood,
// the staging buffers of the production amplitudes and results, and the fixed
// launch geometry that makes every evaluation's reductions deterministic.
// It replaces the static singleton state of cuda::likelihoodInterface
// (cuda/likelihoodInterface.cu:94-107): the amplitudes are uploaded once
// (loadDecayAmps, likelihoodInterface.cu:195-241), nothing is allocated per
// evaluation (the reference allocates and frees device memory in every
// logLikelihood / logLikelihoodDeriv call, likelihoodInterface.cu:378-415),
// and the host waits once per evaluation instead of after every kernel.
#include "internal.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

namespace rpwa_amd {

static thread_local std::string g_lastError;

void set_error(const std::string& msg) { g_lastError = msg; }
void clear_error() { g_lastError.clear(); }

namespace {

int ilog2(int v)
{
	int r = 0;
	while ((1 << r) < v)
		++r;
	return r;
}

uint32_t sector_slots(uint32_t nw, int lanes, uint32_t perSlot)
{
	const uint32_t per = (uint32_t)lanes * perSlot;
	return (nw + per - 1) / per;
}

// Tile with `lanes` lanes per event: the fewest slots per lane such that the
// two sectors, each on its own lanes, fit into the group; false if more than
// 8 slots would be needed.
bool make_tile(const AmpLayout& lay, int lanes, TileShape& ts)
{
	for (int k = 1; k <= 8; ++k) {
		const uint32_t l0 = sector_slots(lay.nW[0], k, lay.perSlot);
		const uint32_t l1 = sector_slots(lay.nW[1], k, lay.perSlot);
		if (l0 + l1 <= (uint32_t)lanes) {
			ts.lanes = lanes;
			ts.slots = k;
			ts.lanes0 = (int)l0;
			ts.lanes1 = (int)l1;
			return true;
		}
	}
	return false;
}

// Tile of the LDS-staged kernel: the largest feasible number of events per
// tile, at most maxEvents; gm = 1 asks for the gradient phase on FP32 MFMA.
bool make_stage_tile(const AmpLayout& lay, int rank, TileShape& ts, int maxEvents = 64, int gm = 0)
{
	for (int e : {64, 32, 16}) {
		StagePlan p;
		if (e <= maxEvents && stage_plan(lay, e, rank, p, gm)) {
			ts = TileShape();
			ts.lanes = 1;
			ts.slots = e;
			ts.events = e;
			ts.gm = gm;
			return true;
		}
	}
	return false;
}

// T of every amplitude-row column for the staged kernel: [rowLen][ranks],
// zero on the padding columns and for ranks beyond `rank`.
void fill_tpad(const AmpLayout& lay, const double2* prod, int rank, int ranks, double2* tpad)
{
	for (uint32_t col = 0; col < lay.rowLen; ++col) {
		const int s = col >= lay.secOff[1] ? 1 : 0;
		const uint32_t i = col - lay.secOff[s];
		const bool valid = i < lay.nW[s];
		for (int r = 0; r < ranks; ++r)
			tpad[(size_t)col * ranks + r] =
			    valid && r < rank ? prod[((size_t)r * 2 + s) * lay.wmax + i] : make_double2(0., 0.);
	}
}

// Modelled default tile (used when autotuning is off): VALU instructions per
// event of the complex MACs of the group's slots, the per-event log/division
// and the butterflies, rank 2 assumed.
bool choose_tile(const AmpLayout& lay, TileShape& best)
{
	double bestCost = 1e300;
	bool found = false;
	for (int lanes : {8, 16, 32, 64}) {
		TileShape ts;
		if (!make_tile(lay, lanes, ts))
			continue;
		const double perWave = ts.slots * (double)lay.perSlot * 2 * 8 + 60 + 8.0 * 3 * ilog2(lanes);
		const double cost = perWave * lanes / 64.0;
		if (cost < bestCost) {
			bestCost = cost;
			best = ts;
			found = true;
		}
	}
	return found;
}

int check_ctx(rpwa_lik_ctx* ctx, const char* fn)
{
	if (!ctx) {
		set_error(std::string(fn) + ": null context");
		return RPWA_ERR_INVALID;
	}
	return RPWA_OK;
}

void free_buffers(rpwa_lik_ctx* c)
{
	hipFree(c->d_amps);
	hipFree(c->d_prod);
	hipFree(c->d_partials);
	hipFree(c->d_sigma);
	hipHostFree(c->h_prod);
	hipHostFree(c->h_out);
	c->h_flag = nullptr;
	c->d_flag = nullptr;
	c->d_amps = nullptr;
	c->d_sigma = nullptr;
	c->rankCap = 0;
	c->partialCap = 0;
	c->d_prod = nullptr;
	c->d_partials = nullptr;
	c->d_out = nullptr;
	c->h_prod = nullptr;
	c->h_out = nullptr;
	c->loaded = false;
}

// (Re)compute the persistent grid of a rank slot's tile (the high-rank slot
// runs the staged kernel instantiated for kMaxRank ranks).
int configure_grid(rpwa_lik_ctx* c, int slot)
{
	int wgPerCu = 1;
	int st = event_kernel_occupancy(c->lay, c->tile[slot], std::min(slot, kMaxRank), &wgPerCu);
	if (st != RPWA_OK) {
		set_error("hipOccupancyMaxActiveBlocksPerMultiprocessor failed");
		return st;
	}
	c->nWg[slot] = c->nCu * std::min(wgPerCu, 8);
	return RPWA_OK;
}

// Buffers of the evaluation sized for `rank` (at least kMaxRank): production
// amplitudes and the staged kernel's T per row column for every rank block
// (device and pinned host), the result row, the workgroup rows of the largest
// grid any tile can get (8 workgroups per CU), and for ranks above kMaxRank the
// per-event sigma.  Called only while no evaluation is in flight.
int ensure_rank_buffers(rpwa_lik_ctx* c, uint32_t rank)
{
	const uint32_t r = std::max<uint32_t>(rank, (uint32_t)kMaxRank);
	if (r > (uint32_t)kMaxRank && !c->d_sigma)
		RPWA_HIP_CHECK(hipMalloc(&c->d_sigma, std::max<size_t>(c->lay.nEvents, 1) * sizeof(double)));
	if (r <= c->rankCap)
		return RPWA_OK;
	const size_t blocks = (r + kMaxRank - 1) / kMaxRank;
	const size_t nProd = (size_t)r * 2 * c->lay.wmax + blocks * c->lay.rowLen * kMaxRank;
	const size_t nOut = partial_elems((int)r, c->lay.wmax);
	const size_t need = (size_t)c->nCu * 8 * nOut;
	hipFree(c->d_prod);
	hipFree(c->d_partials);
	hipHostFree(c->h_prod);
	hipHostFree(c->h_out);
	c->d_prod = nullptr;
	c->d_partials = nullptr;
	c->h_prod = nullptr;
	c->h_out = nullptr;
	c->d_out = nullptr;
	c->rankCap = 0;
	c->partialCap = 0;
	RPWA_HIP_CHECK(hipMalloc(&c->d_prod, nProd * sizeof(double2)));
	RPWA_HIP_CHECK(hipHostMalloc(&c->h_prod, nProd * sizeof(double2), hipHostMallocDefault));
	// results land in coherent pinned memory written by the reduce kernel itself (no copy-back command),
	// followed by one 8-byte completion flag per block of the reduction
	const size_t nFlags = reduce_blocks((uint32_t)nOut);
	RPWA_HIP_CHECK(hipHostMalloc(&c->h_out, (nOut + nFlags) * sizeof(double),
	                             hipHostMallocMapped | hipHostMallocCoherent));
	RPWA_HIP_CHECK(hipHostGetDevicePointer(reinterpret_cast<void**>(&c->d_out), c->h_out, 0));
	c->h_flag = reinterpret_cast<unsigned long long*>(c->h_out + nOut);
	c->d_flag = reinterpret_cast<unsigned long long*>(c->d_out + nOut);
	for (size_t b = 0; b < nFlags; ++b)
		c->h_flag[b] = 0;
	c->seq = 0;

	RPWA_HIP_CHECK(hipMalloc(&c->d_partials, need * sizeof(double)));
	c->partialCap = need;
	c->rankCap = r;
	return RPWA_OK;
}

int setup_layout(rpwa_lik_ctx* c, uint64_t nEvents, const uint32_t nW[2])
{
	AmpLayout& L = c->lay;
	L.nEvents = nEvents;
	L.nW[0] = nW[0];
	L.nW[1] = nW[1];
	L.wmax = std::max(nW[0], nW[1]);
	L.precision = c->precision;
	L.perSlot = c->precision == RPWA_PREC_MIXED ? 2 : 1;
	const uint32_t pad0 = (nW[0] + L.perSlot - 1) / L.perSlot * L.perSlot;
	const uint32_t pad1 = (nW[1] + L.perSlot - 1) / L.perSlot * L.perSlot;
	L.secOff[0] = 0;
	L.secOff[1] = pad0;
	L.rowLen = pad0 + pad1;
	TileShape ts;
	if (!choose_tile(L, ts)) {
		set_error("rpwa_lik: " + std::to_string(nW[0]) + "+" + std::to_string(nW[1]) +
		          " waves exceed the register tiles (at most " + std::to_string(8 * 64 * L.perSlot) +
		          " waves in total)");
		return RPWA_ERR_UNSUPPORTED;
	}
	for (int r = 0; r <= kHighRankSlot; ++r) {
		c->tile[r] = ts;
		c->tileFixed[r] = false;
		c->evaluated[r] = false;
	}
	return RPWA_OK;
}

int alloc_eval_buffers(rpwa_lik_ctx* c)
{
	if (int st = ensure_rank_buffers(c, (uint32_t)kMaxRank))
		return st;
	for (int r = 1; r <= kMaxRank; ++r)
		if (int st = configure_grid(c, r))
			return st;
	return RPWA_OK;
}

// load from FP64 amplitudes already on the device
int load_from_device(rpwa_lik_ctx* c, const double* d_src, int srcLayout, uint64_t nEvents,
                     const uint32_t nW[2], const double* waveNorm)
{
	free_buffers(c);
	int st = setup_layout(c, nEvents, nW);
	if (st != RPWA_OK)
		return st;
	RPWA_HIP_CHECK(hipMalloc(&c->d_amps, std::max<size_t>(c->lay.bytes(), 16)));
	double* d_norm = nullptr;
	if (waveNorm) {
		const uint32_t W = nW[0] + nW[1];
		RPWA_HIP_CHECK(hipMalloc(&d_norm, W * sizeof(double)));
		RPWA_HIP_CHECK(hipMemcpy(d_norm, waveNorm, W * sizeof(double), hipMemcpyHostToDevice));
	}
	hipError_t e = launch_transform_amps(c->lay, d_src, srcLayout, d_norm, c->d_amps, c->stream);
	if (e == hipSuccess)
		e = hipStreamSynchronize(c->stream);
	hipFree(d_norm);
	RPWA_HIP_CHECK(e);
	st = alloc_eval_buffers(c);
	if (st != RPWA_OK)
		return st;
	c->loaded = true;
	return RPWA_OK;
}

// Measured tile choice for one rank: every feasible lane count
// runs the event kernel a few times on the context's amplitudes with fixed
// production amplitudes; the fastest (median of the timed launches) is kept.
int autotune_rank(rpwa_lik_ctx* c, int rank)
{
	const uint32_t wmax = c->lay.wmax;
	const size_t nProd = (size_t)rank * 2 * wmax;
	const int ranks = stage_rank_cap(rank);
	std::vector<double2> prod(nProd + (size_t)c->lay.rowLen * ranks);
	for (size_t i = 0; i < nProd; ++i)
		prod[i] = make_double2(0.5 + 0.01 * (double)(i % 7), 0.25 - 0.01 * (double)(i % 5));
	fill_tpad(c->lay, prod.data(), rank, ranks, prod.data() + nProd);
	RPWA_HIP_CHECK(hipMemcpyAsync(c->d_prod, prod.data(), prod.size() * sizeof(double2), hipMemcpyHostToDevice,
	                              c->stream));
	const uint32_t nElem = partial_elems(rank, wmax);
	EventArgs a;
	a.amps = c->d_amps;
	a.nEvents = c->lay.nEvents;
	a.rowLen = c->lay.rowLen;
	a.secOff0 = c->lay.secOff[0];
	a.secOff1 = c->lay.secOff[1];
	a.nW0 = c->lay.nW[0];
	a.nW1 = c->lay.nW[1];
	a.wmax = wmax;
	a.rank = rank;
	a.prod = reinterpret_cast<const double2*>(c->d_prod);
	a.tpad = reinterpret_cast<const double2*>(c->d_prod) + nProd;
	a.tpadHost = prod.data() + nProd;
	a.flat2 = 0.01;
	a.partials = c->d_partials;
	a.nElem = nElem;
	const TileShape keep = c->tile[rank];
	TileShape best = keep;
	double bestT = 1e300;
	const int reps = 5;
	// register tiles of 8..64 lanes per event, then the LDS-staged kernel with
	// tiles of 64, 32 and 16 events (the smaller tiles give a short event list
	// more tiles per workgroup)
	// (and, for complex64 rows of at most 64 units, the staged kernel with the
	// gradient phase on FP32 MFMA, tiles of 64 and 32 events)
	// (and the wave-tile variant: S and the gradient on MFMA, one wave per 16 events, 64-event tiles)
	const int cands[][2] = {{8, 0}, {16, 0}, {32, 0}, {64, 0}, {-64, 0}, {-32, 0}, {-16, 0}, {-64, 1}, {-32, 1},
	                        {-64, 2}};
	for (const auto& cg : cands) {
			const int cand = cg[0], gm = cg[1];
			TileShape ts;
			if (cand > 0 ? !make_tile(c->lay, cand, ts)
			             : (!make_stage_tile(c->lay, rank, ts, -cand, gm) || ts.events != -cand || ts.gm != gm))
				continue;
			c->tile[rank] = ts;
			if (int st = configure_grid(c, rank))
				return st;
			a.partials = c->d_partials;  // the grid configuration owns the buffer
			a.lanes0 = ts.lanes0;
			a.lanes1 = ts.lanes1;
			std::vector<float> ms;
			for (int it = 0; it < reps + 1; ++it) {
				// the reduction of the workgroup rows is timed with the kernel: its cost
				// grows with the grid, which differs between the variants
				RPWA_HIP_CHECK(launch_event_kernel(c->lay, ts, a, true, c->nWg[rank], c->stream, c->ev0, nullptr));
				RPWA_HIP_CHECK(launch_event_reduce(c->d_partials, c->nWg[rank], nElem, 0.1, c->d_out, c->stream,
				                                   nullptr, 0, nullptr, c->ev1));
				RPWA_HIP_CHECK(hipEventSynchronize(c->ev1));
				float t = 0.f;
				RPWA_HIP_CHECK(hipEventElapsedTime(&t, c->ev0, c->ev1));
				if (it > 0)
					ms.push_back(t);
			}
			std::sort(ms.begin(), ms.end());
			const double med = ms[ms.size() / 2];
			if (std::getenv("RPWA_VERBOSE"))
				std::fprintf(stderr, "[rpwa autotune] precision %d rank %d lanes %d slots %d (%d+%d lanes) "
				             "events/tile %d%s workgroups %d: %.4f ms\n", c->precision, rank, ts.lanes, ts.slots,
				             ts.lanes0, ts.lanes1, ts.events,
				             ts.gm == 2 ? " (S and gradient on MFMA)" : ts.gm ? " (gradient on MFMA)" : "",
				             c->nWg[rank], med);
			if (med < bestT) {
				bestT = med;
				best = ts;
			}
		}
	c->tile[rank] = best;
	return configure_grid(c, rank);
}

// Start one evaluation on the context's stream; results are collected by
// collect().  The whole result row (value, flat derivative, derivatives) is
// always copied back: a few kB.  Ranks up to kMaxRank: one fused event kernel;
// higher ranks: the staged kernel in rank blocks (sigma of every block, then
// the gradient entries of every block), then the same reduction.
int submit(rpwa_lik_ctx* c, const double* prodAmps, uint32_t nmbProdAmps, double flat, uint32_t rank,
           const char* fn)
{
	if (!c->loaded) {
		set_error(std::string(fn) + ": decay amplitudes not loaded");
		return RPWA_ERR_STATE;
	}
	if (c->pending) {
		set_error(std::string(fn) + ": an evaluation is still pending (rpwa_lik_collect first)");
		return RPWA_ERR_STATE;
	}
	if (!prodAmps || rank == 0) {
		set_error(std::string(fn) + ": production amplitudes missing or rank 0");
		return RPWA_ERR_INVALID;
	}
	const uint32_t wmax = c->lay.wmax;
	if ((uint64_t)nmbProdAmps != (uint64_t)rank * 2 * wmax) {
		set_error(std::string(fn) + ": nmbProdAmps must be rank*2*nmbWavesMax = " +
		          std::to_string((uint64_t)rank * 2 * wmax));
		return RPWA_ERR_INVALID;
	}
	RPWA_HIP_CHECK(hipSetDevice(c->device));
	if (int st = ensure_rank_buffers(c, rank))
		return st;
	const int slot = rank_slot(rank);
	if (!c->tileFixed[slot]) {
		if (slot == kHighRankSlot) {
			// ranks above kMaxRank: the staged kernel with the largest tile that fits
			TileShape ts;
			if (!make_stage_tile(c->lay, kMaxRank, ts)) {
				set_error(std::string(fn) + ": rank " + std::to_string(rank) + " needs the LDS-staged kernel, "
				          "and the amplitude row does not fit it");
				return RPWA_ERR_UNSUPPORTED;
			}
			c->tile[slot] = ts;
			if (int st = configure_grid(c, slot))
				return st;
		} else if (c->autotune) {
			if (int st = autotune_rank(c, (int)rank))
				return st;
		}
		c->tileFixed[slot] = true;
	}
	std::memcpy(c->h_prod, prodAmps, (size_t)nmbProdAmps * sizeof(double2));
	const bool blocks = rank > (uint32_t)kMaxRank;
	const bool staged = c->tile[slot].events > 0;
	const uint32_t nBlocks = blocks ? (rank + kMaxRank - 1) / kMaxRank : 1;
	const int ranks = blocks ? kMaxRank : stage_rank_cap((int)rank);
	size_t nUp = nmbProdAmps;
	if (staged) {
		// T per row column, one [rowLen][ranks] array per rank block
		for (uint32_t b = 0; b < nBlocks; ++b) {
			const uint32_t r0 = b * kMaxRank;
			const int nr = (int)std::min<uint32_t>(rank - r0, (uint32_t)ranks);
			fill_tpad(c->lay, reinterpret_cast<const double2*>(c->h_prod) + (size_t)r0 * 2 * wmax, nr, ranks,
			          reinterpret_cast<double2*>(c->h_prod) + nUp);
			nUp += (size_t)c->lay.rowLen * ranks;
		}
	}
	// the staged kernel takes T in its arguments when it fits: nothing to upload
	if (!(staged && !blocks && stage_t_inline(c->lay, (int)rank)))
		RPWA_HIP_CHECK(hipMemcpyAsync(c->d_prod, c->h_prod, nUp * sizeof(double2), hipMemcpyHostToDevice,
		                              c->stream));
	const uint32_t nElem = partial_elems((int)rank, wmax);
	EventArgs a;
	a.amps = c->d_amps;
	a.nEvents = c->lay.nEvents;
	a.rowLen = c->lay.rowLen;
	a.secOff0 = c->lay.secOff[0];
	a.secOff1 = c->lay.secOff[1];
	a.nW0 = c->lay.nW[0];
	a.nW1 = c->lay.nW[1];
	a.wmax = wmax;
	a.rank = blocks ? kMaxRank : (int)rank;
	a.lanes0 = c->tile[slot].lanes0;
	a.lanes1 = c->tile[slot].lanes1;
	a.prod = reinterpret_cast<const double2*>(c->d_prod);
	a.tpad = reinterpret_cast<const double2*>(c->d_prod) + nmbProdAmps;
	a.tpadHost = blocks ? nullptr : reinterpret_cast<const double2*>(c->h_prod) + nmbProdAmps;
	a.flat2 = flat * flat;
	a.partials = c->d_partials;
	a.nElem = nElem;
	a.sigma = c->d_sigma;
	// timing events are bound to the launches: the first event launch starts ev0,
	// the last one stops evMid, the reduction runs between evRed and ev1
	const bool timed = c->timing != 0 && (c->seq + 1) % c->timing == 0;
	hipEvent_t kStart = timed ? c->ev0 : nullptr, kStop = timed ? c->evMid : nullptr;
	if (!blocks) {
		RPWA_HIP_CHECK(launch_event_kernel(c->lay, c->tile[slot], a, true, c->nWg[slot], c->stream, kStart, kStop));
	} else {
		const size_t tStride = (size_t)c->lay.rowLen * kMaxRank;
		for (int pass = 0; pass < 2; ++pass)
			for (uint32_t b = 0; b < nBlocks; ++b) {
				a.stageMode = pass == 0 ? STAGE_SIGMA : STAGE_GRAD;
				a.accumulate = b > 0 ? 1u : 0u;
				a.rank0 = b * kMaxRank;
				a.tpad = reinterpret_cast<const double2*>(c->d_prod) + nmbProdAmps + b * tStride;
				const bool first = pass == 0 && b == 0, last = pass == 1 && b + 1 == nBlocks;
				RPWA_HIP_CHECK(launch_stage_kernel(c->lay, c->tile[slot].events, c->tile[slot].gm, a, c->nWg[slot], c->stream,
				                                   first ? kStart : nullptr, last ? kStop : nullptr));
			}
	}
	const unsigned long long seq = ++c->seq;
	RPWA_HIP_CHECK(launch_event_reduce(c->d_partials, c->nWg[slot], nElem, flat, c->d_out, c->stream,
	                                   c->d_flag, seq, timed ? c->evRed : nullptr,
	                                   timed ? c->ev1 : nullptr));
	c->pendingBlocks = reduce_blocks(nElem);
	c->pendingTimed = timed;
	c->pending = true;
	c->evaluated[slot] = true;
	c->pendingRank = rank;
	c->pendingProd = nmbProdAmps;
	return RPWA_OK;
}

int collect(rpwa_lik_ctx* c, double* lnl, double* deriv, double* derivFlat, const char* fn)
{
	if (!c->pending) {
		set_error(std::string(fn) + ": no evaluation pending");
		return RPWA_ERR_STATE;
	}
	c->pending = false;
	RPWA_HIP_CHECK(hipSetDevice(c->device));
	// Every block of the reduction stores the evaluation's sequence number into
	// its mapped flag after its results: watching them avoids the wake-up latency
	// of a blocking synchronisation.  Flags that do not arrive within a second (a
	// failed launch, a long kernel) fall back to the stream synchronisation,
	// which also reports errors.
	bool done = false;
	{
		const volatile unsigned long long* flag = c->h_flag;
		const auto t0 = std::chrono::steady_clock::now();
		uint32_t seen = 0;  // blocks 0..seen-1 are done
		for (uint32_t spin = 0;; ++spin) {
			while (seen < c->pendingBlocks && flag[seen] == c->seq)
				++seen;
			if (seen == c->pendingBlocks) {
				done = true;
				break;
			}
			if ((spin & 1023u) == 1023u && std::chrono::steady_clock::now() - t0 > std::chrono::seconds(1))
				break;
			__builtin_ia32_pause();
		}
		std::atomic_thread_fence(std::memory_order_acquire);
	}
	if (!done)
		RPWA_HIP_CHECK(hipStreamSynchronize(c->stream));
	// the timing events complete with the reduction's dispatch, just after its
	// flags: waiting for the last one costs little once the flags are seen
	if (c->pendingTimed)
		RPWA_HIP_CHECK(hipEventSynchronize(c->ev1));
	float ms = 0.f, ms1 = 0.f;
	if (c->pendingTimed && hipEventElapsedTime(&ms, c->ev0, c->evMid) == hipSuccess &&
	    hipEventElapsedTime(&ms1, c->evRed, c->ev1) == hipSuccess) {
		c->eventTime += 1e-3 * ms;
		c->reduceTime += 1e-3 * ms1;
		c->kernelTime += 1e-3 * ((double)ms + (double)ms1);
		++c->nEvals;
	}
	if (lnl)
		*lnl = c->h_out[0];
	if (derivFlat)
		*derivFlat = c->h_out[1];
	if (deriv)
		std::memcpy(deriv, c->h_out + 2, (size_t)c->pendingProd * sizeof(double2));
	return RPWA_OK;
}

int evaluate(rpwa_lik_ctx* c, const double* prodAmps, uint32_t nmbProdAmps, double flat, uint32_t rank,
             double* lnl, double* deriv, double* derivFlat, const char* fn)
{
	if (int st = submit(c, prodAmps, nmbProdAmps, flat, rank, fn))
		return st;
	return collect(c, lnl, deriv, derivFlat, fn);
}

}  // namespace
}  // namespace rpwa_amd

using namespace rpwa_amd;

#define RPWA_GUARD_BEGIN try {
#define RPWA_GUARD_END                                  \
	}                                                   \
	catch (const std::bad_alloc&) {                     \
		set_error("host allocation failed");            \
		return RPWA_ERR_NOMEM;                          \
	}                                                   \
	catch (const std::exception& ex) {                  \
		set_error(ex.what());                           \
		return RPWA_ERR_INVALID;                        \
	}

extern "C" const char* rpwa_last_error(void) { return g_lastError.c_str(); }

extern "C" int rpwa_device_count(void)
{
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess)
		return 0;
	return n;
}

extern "C" int rpwa_lik_create(rpwa_lik_ctx** ctx, int device, int precision)
{
	clear_error();
	if (!ctx) {
		set_error("rpwa_lik_create: null output pointer");
		return RPWA_ERR_INVALID;
	}
	*ctx = nullptr;
	if (precision != RPWA_PREC_FP64 && precision != RPWA_PREC_MIXED) {
		set_error("rpwa_lik_create: unknown precision");
		return RPWA_ERR_INVALID;
	}
	const int n = rpwa_device_count();
	if (device < 0 || device >= n) {
		set_error("rpwa_lik_create: no GPU with index " + std::to_string(device) + " (" +
		          std::to_string(n) + " visible)");
		return RPWA_ERR_INVALID;
	}
	RPWA_GUARD_BEGIN
	RPWA_HIP_CHECK(hipSetDevice(device));
	rpwa_lik_ctx* c = new rpwa_lik_ctx();
	c->device = device;
	c->precision = precision;
	hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
	if (e == hipSuccess)
		e = hipEventCreate(&c->ev0);
	if (e == hipSuccess)
		e = hipEventCreate(&c->ev1);
	if (e == hipSuccess)
		e = hipEventCreate(&c->evMid);
	if (e == hipSuccess)
		e = hipEventCreate(&c->evRed);
	if (e == hipSuccess)
		e = hipDeviceGetAttribute(&c->nCu, hipDeviceAttributeMultiprocessorCount, device);
	if (const char* env = std::getenv("RPWA_AUTOTUNE"))
		c->autotune = std::string(env) != "0";
	if (const char* env = std::getenv("RPWA_KERNEL_TIMING"))
		c->timing = (uint32_t)std::max(0, std::atoi(env));
	if (e != hipSuccess) {
		set_error(std::string("rpwa_lik_create: ") + hipGetErrorString(e));
		for (hipEvent_t ev : {c->ev0, c->ev1, c->evMid, c->evRed})
			if (ev)
				hipEventDestroy(ev);
		if (c->stream)
			hipStreamDestroy(c->stream);
		delete c;
		return RPWA_ERR_HIP;
	}
	*ctx = c;
	return RPWA_OK;
	RPWA_GUARD_END
}

extern "C" int rpwa_lik_retain(rpwa_lik_ctx* ctx)
{
	clear_error();
	if (int st = check_ctx(ctx, "rpwa_lik_retain"))
		return st;
	ctx->refs.fetch_add(1);
	return RPWA_OK;
}

extern "C" int rpwa_lik_release(rpwa_lik_ctx* ctx)
{
	clear_error();
	if (int st = check_ctx(ctx, "rpwa_lik_release"))
		return st;
	if (ctx->refs.fetch_sub(1) == 1) {
		hipSetDevice(ctx->device);
		hipStreamSynchronize(ctx->stream);
		free_buffers(ctx);
		hipEventDestroy(ctx->ev0);
		hipEventDestroy(ctx->ev1);
		hipEventDestroy(ctx->evMid);
		hipEventDestroy(ctx->evRed);
		hipStreamDestroy(ctx->stream);
		delete ctx;
	}
	return RPWA_OK;
}

extern "C" int rpwa_lik_load_decay_amps(rpwa_lik_ctx* ctx, const double* decayAmps, uint64_t nmbDecayAmps,
                                        uint64_t nmbEvents, const uint32_t nmbWavesRefl[2], int layout)
{
	clear_error();
	if (int st = check_ctx(ctx, "rpwa_lik_load_decay_amps"))
		return st;
	if (!decayAmps || !nmbWavesRefl || nmbEvents == 0 || nmbWavesRefl[0] + nmbWavesRefl[1] == 0) {
		set_error("rpwa_lik_load_decay_amps: empty or null input");
		return RPWA_ERR_INVALID;
	}
	const uint64_t W = (uint64_t)nmbWavesRefl[0] + nmbWavesRefl[1];
	const uint64_t wmax = std::max(nmbWavesRefl[0], nmbWavesRefl[1]);
	uint64_t expect = 0;
	if (layout == RPWA_LAYOUT_EVT_WAVE)
		expect = nmbEvents * W;
	else if (layout == RPWA_LAYOUT_REFL_WAVE_EVT)
		expect = 2 * wmax * nmbEvents;
	else {
		set_error("rpwa_lik_load_decay_amps: unknown layout");
		return RPWA_ERR_INVALID;
	}
	if (nmbDecayAmps != expect) {
		set_error("rpwa_lik_load_decay_amps: nmbDecayAmps = " + std::to_string(nmbDecayAmps) +
		          " does not match the layout (" + std::to_string(expect) + ")");
		return RPWA_ERR_INVALID;
	}
	RPWA_GUARD_BEGIN
	RPWA_HIP_CHECK(hipSetDevice(ctx->device));
	double* d_src = nullptr;
	RPWA_HIP_CHECK(hipMalloc(&d_src, expect * sizeof(double2)));
	hipError_t e = hipMemcpy(d_src, decayAmps, expect * sizeof(double2), hipMemcpyHostToDevice);
	if (e != hipSuccess) {
		hipFree(d_src);
		RPWA_HIP_CHECK(e);
	}
	const int st = load_from_device(ctx, d_src, layout, nmbEvents, nmbWavesRefl, nullptr);
	hipFree(d_src);
	return st;
	RPWA_GUARD_END
}

extern "C" int rpwa_lik_load_decay_amps_device(rpwa_lik_ctx* ctx, const double* d_amps, uint64_t nmbEvents,
                                               const uint32_t nmbWavesRefl[2], const double* waveNorm)
{
	clear_error();
	if (int st = check_ctx(ctx, "rpwa_lik_load_decay_amps_device"))
		return st;
	if (!d_amps || !nmbWavesRefl || nmbEvents == 0 || nmbWavesRefl[0] + nmbWavesRefl[1] == 0) {
		set_error("rpwa_lik_load_decay_amps_device: empty or null input");
		return RPWA_ERR_INVALID;
	}
	RPWA_GUARD_BEGIN
	RPWA_HIP_CHECK(hipSetDevice(ctx->device));
	return load_from_device(ctx, d_amps, RPWA_LAYOUT_EVT_WAVE, nmbEvents, nmbWavesRefl, waveNorm);
	RPWA_GUARD_END
}

extern "C" int rpwa_lik_log_likelihood(rpwa_lik_ctx* ctx, const double* prodAmps, uint32_t nmbProdAmps,
                                       double prodAmpFlat, uint32_t rank, double* logLikelihood)
{
	clear_error();
	if (int st = check_ctx(ctx, "rpwa_lik_log_likelihood"))
		return st;
	if (!logLikelihood) {
		set_error("rpwa_lik_log_likelihood: null output");
		return RPWA_ERR_INVALID;
	}
	RPWA_GUARD_BEGIN
	return evaluate(ctx, prodAmps, nmbProdAmps, prodAmpFlat, rank, logLikelihood, nullptr, nullptr,
	                "rpwa_lik_log_likelihood");
	RPWA_GUARD_END
}

extern "C" int rpwa_lik_log_likelihood_deriv(rpwa_lik_ctx* ctx, const double* prodAmps, uint32_t nmbProdAmps,
                                             double prodAmpFlat, uint32_t rank, double* derivatives,
                                             double* derivativeFlat)
{
	clear_error();
	if (int st = check_ctx(ctx, "rpwa_lik_log_likelihood_deriv"))
		return st;
	if (!derivatives || !derivativeFlat) {
		set_error("rpwa_lik_log_likelihood_deriv: null output");
		return RPWA_ERR_INVALID;
	}
	RPWA_GUARD_BEGIN
	return evaluate(ctx, prodAmps, nmbProdAmps, prodAmpFlat, rank, nullptr, derivatives, derivativeFlat,
	                "rpwa_lik_log_likelihood_deriv");
	RPWA_GUARD_END
}

extern "C" int rpwa_lik_fdf(rpwa_lik_ctx* ctx, const double* prodAmps, uint32_t nmbProdAmps, double prodAmpFlat,
                            uint32_t rank, double* logLikelihood, double* derivatives, double* derivativeFlat)
{
	clear_error();
	if (int st = check_ctx(ctx, "rpwa_lik_fdf"))
		return st;
	if (!logLikelihood || !derivatives || !derivativeFlat) {
		set_error("rpwa_lik_fdf: null output");
		return RPWA_ERR_INVALID;
	}
	RPWA_GUARD_BEGIN
	return evaluate(ctx, prodAmps, nmbProdAmps, prodAmpFlat, rank, logLikelihood, derivatives,
	                derivativeFlat, "rpwa_lik_fdf");
	RPWA_GUARD_END
}

extern "C" int rpwa_lik_submit(rpwa_lik_ctx* ctx, const double* prodAmps, uint32_t nmbProdAmps,
                               double prodAmpFlat, uint32_t rank)
{
	clear_error();
	if (int st = check_ctx(ctx, "rpwa_lik_submit"))
		return st;
	RPWA_GUARD_BEGIN
	return submit(ctx, prodAmps, nmbProdAmps, prodAmpFlat, rank, "rpwa_lik_submit");
	RPWA_GUARD_END
}

extern "C" int rpwa_lik_collect(rpwa_lik_ctx* ctx, double* logLikelihood, double* derivatives,
                                double* derivativeFlat)
{
	clear_error();
	if (int st = check_ctx(ctx, "rpwa_lik_collect"))
		return st;
	RPWA_GUARD_BEGIN
	return collect(ctx, logLikelihood, derivatives, derivativeFlat, "rpwa_lik_collect");
	RPWA_GUARD_END
}

extern "C" int rpwa_lik_kernel_time(rpwa_lik_ctx* ctx, double* seconds)
{
	clear_error();
	if (int st = check_ctx(ctx, "rpwa_lik_kernel_time"))
		return st;
	if (!seconds) {
		set_error("rpwa_lik_kernel_time: null output");
		return RPWA_ERR_INVALID;
	}
	*seconds = ctx->kernelTime;
	return RPWA_OK;
}

extern "C" int rpwa_lik_reset_kernel_time(rpwa_lik_ctx* ctx)
{
	clear_error();
	if (int st = check_ctx(ctx, "rpwa_lik_reset_kernel_time"))
		return st;
	ctx->kernelTime = 0;
	ctx->eventTime = 0;
	ctx->reduceTime = 0;
	ctx->nEvals = 0;
	return RPWA_OK;
}

extern "C" int rpwa_lik_kernel_times(rpwa_lik_ctx* ctx, double* eventKernel, double* reduceKernel,
                                     uint64_t* nmbEvaluations)
{
	clear_error();
	if (int st = check_ctx(ctx, "rpwa_lik_kernel_times"))
		return st;
	if (eventKernel)
		*eventKernel = ctx->eventTime;
	if (reduceKernel)
		*reduceKernel = ctx->reduceTime;
	if (nmbEvaluations)
		*nmbEvaluations = ctx->nEvals;
	return RPWA_OK;
}

extern "C" int rpwa_lik_info(rpwa_lik_ctx* ctx, uint64_t info[9])
{
	clear_error();
	if (int st = check_ctx(ctx, "rpwa_lik_info"))
		return st;
	if (!info) {
		set_error("rpwa_lik_info: null output");
		return RPWA_ERR_INVALID;
	}
	info[0] = ctx->lay.nEvents;
	info[1] = ctx->lay.nW[0];
	info[2] = ctx->lay.nW[1];
	info[3] = (uint64_t)ctx->precision;
	int r = 1;
	for (int k = kHighRankSlot; k >= 1; --k)
		if (ctx->evaluated[k])
			r = k;  // lowest rank evaluated so far (1 if none)
	info[4] = (uint64_t)ctx->tile[r].lanes;
	info[5] = (uint64_t)ctx->tile[r].slots;
	info[6] = (uint64_t)ctx->nWg[r];
	info[7] = 256;
	info[8] = ctx->lay.bytes();
	return RPWA_OK;
}

extern "C" int rpwa_lik_set_tile(rpwa_lik_ctx* ctx, uint32_t lanesPerEvent)
{
	clear_error();
	if (int st = check_ctx(ctx, "rpwa_lik_set_tile"))
		return st;
	if (!ctx->loaded) {
		set_error("rpwa_lik_set_tile: decay amplitudes not loaded");
		return RPWA_ERR_STATE;
	}
	const uint32_t lanes = lanesPerEvent;
	if (lanes != 1 && lanes != 8 && lanes != 16 && lanes != 32 && lanes != 64) {
		set_error("rpwa_lik_set_tile: lanes per event must be 1 (LDS-staged), 8, 16, 32 or 64");
		return RPWA_ERR_INVALID;
	}
	TileShape ts;
	if (lanes == 1 ? !make_stage_tile(ctx->lay, kMaxRank, ts) : !make_tile(ctx->lay, (int)lanes, ts)) {
		set_error(lanes == 1 ? std::string("rpwa_lik_set_tile: the amplitude row does not fit the LDS-staged "
		                                   "kernel (at most 256 16-byte units per row)")
		                     : "rpwa_lik_set_tile: the amplitude row does not fit " + std::to_string(lanes) +
		                           " lanes x 8 slots");
		return RPWA_ERR_UNSUPPORTED;
	}
	RPWA_GUARD_BEGIN
	RPWA_HIP_CHECK(hipSetDevice(ctx->device));
	for (int r = 1; r <= kMaxRank; ++r) {
		if (lanes == 1)
			make_stage_tile(ctx->lay, r, ts);
		ctx->tile[r] = ts;
		ctx->tileFixed[r] = true;
		if (int st = configure_grid(ctx, r))
			return st;
	}
	// ranks above kMaxRank always run the staged kernel in rank blocks
	TileShape high;
	ctx->tileFixed[kHighRankSlot] = false;
	if (make_stage_tile(ctx->lay, kMaxRank, high)) {
		ctx->tile[kHighRankSlot] = high;
		ctx->tileFixed[kHighRankSlot] = true;
		if (int st = configure_grid(ctx, kHighRankSlot))
			return st;
	}
	return RPWA_OK;
	RPWA_GUARD_END
}

namespace {
int set_stage_variant(rpwa_lik_ctx* ctx, uint32_t eventsPerTile, int gm, const char* fn)
{
	clear_error();
	if (int st = check_ctx(ctx, fn))
		return st;
	if (!ctx->loaded) {
		set_error(std::string(fn) + ": decay amplitudes not loaded");
		return RPWA_ERR_STATE;
	}
	const int e = (int)eventsPerTile;
	if (e != 16 && e != 32 && e != 64) {
		set_error(std::string(fn) + ": events per tile must be 16, 32 or 64");
		return RPWA_ERR_INVALID;
	}
	TileShape ts;
	// the wave-tile variant covers ranks 1 and 2; higher ranks keep the VALU staged kernel
	const int rTop = gm >= 2 ? 2 : kMaxRank;
	for (int r = 1; r <= rTop; ++r)
		if (!make_stage_tile(ctx->lay, r, ts, e, gm) || ts.events != e || ts.gm != gm) {
			set_error(std::string(fn) + ": a tile of " + std::to_string(e) + " amplitude rows does not fit the " +
			          (gm == 2 ? "wave-tile variant (complex64 rows of at most 96 waves, ranks up to 2, "
			                     "tiles of 64 events)"
			           : gm ? "LDS-staged kernel with the gradient on MFMA (complex64 rows of at most 128 waves, "
			                  "tiles of 64 or 32 events)"
			                : "LDS-staged kernel"));
			return RPWA_ERR_UNSUPPORTED;
		}
	RPWA_GUARD_BEGIN
	RPWA_HIP_CHECK(hipSetDevice(ctx->device));
	for (int r = 1; r <= kHighRankSlot; ++r) {
		if (!make_stage_tile(ctx->lay, std::min(r, kMaxRank), ts, e, r <= rTop ? gm : 0))
			make_stage_tile(ctx->lay, std::min(r, kMaxRank), ts);
		ctx->tile[r] = ts;
		ctx->tileFixed[r] = true;
		if (int st = configure_grid(ctx, r))
			return st;
	}
	return RPWA_OK;
	RPWA_GUARD_END
}
}  // namespace

extern "C" int rpwa_lik_set_stage_tile(rpwa_lik_ctx* ctx, uint32_t eventsPerTile)
{
	return set_stage_variant(ctx, eventsPerTile, 0, "rpwa_lik_set_stage_tile");
}

extern "C" int rpwa_lik_set_stage_variant(rpwa_lik_ctx* ctx, uint32_t eventsPerTile, int variant)
{
	if (variant < 0 || variant > 2) {
		clear_error();
		set_error("rpwa_lik_set_stage_variant: variant must be 0, 1 or 2");
		return RPWA_ERR_INVALID;
	}
	return set_stage_variant(ctx, eventsPerTile, variant, "rpwa_lik_set_stage_variant");
}

extern "C" int rpwa_lik_kernel_variant(rpwa_lik_ctx* ctx, uint32_t rank, uint32_t variant[4])
{
	clear_error();
	if (int st = check_ctx(ctx, "rpwa_lik_kernel_variant"))
		return st;
	if (!variant || rank == 0) {
		set_error("rpwa_lik_kernel_variant: null output or rank 0");
		return RPWA_ERR_INVALID;
	}
	const int slot = rank_slot(rank);
	variant[0] = (uint32_t)ctx->tile[slot].lanes;
	variant[1] = (uint32_t)ctx->tile[slot].slots;
	variant[2] = (uint32_t)ctx->tile[slot].gm;
	variant[3] = (uint32_t)ctx->nWg[slot];
	return RPWA_OK;
}

extern "C" int rpwa_lik_set_timing(rpwa_lik_ctx* ctx, int enable)
{
	clear_error();
	if (int st = check_ctx(ctx, "rpwa_lik_set_timing"))
		return st;
	if (ctx->pending) {
		set_error("rpwa_lik_set_timing: an evaluation is pending");
		return RPWA_ERR_STATE;
	}
	ctx->timing = enable > 0 ? (uint32_t)enable : 0u;
	return RPWA_OK;
}

extern "C" int rpwa_lik_set_autotune(rpwa_lik_ctx* ctx, int enable)
{
	clear_error();
	if (int st = check_ctx(ctx, "rpwa_lik_set_autotune"))
		return st;
	ctx->autotune = enable != 0;
	return RPWA_OK;
}

extern "C" int rpwa_lik_stream(rpwa_lik_ctx* ctx, void** hipStream)
{
	clear_error();
	if (int st = check_ctx(ctx, "rpwa_lik_stream"))
		return st;
	if (!hipStream) {
		set_error("rpwa_lik_stream: null output");
		return RPWA_ERR_INVALID;
	}
	*hipStream = ctx->stream;
	return RPWA_OK;
}

extern "C" int rpwa_dev_alloc(int device, uint64_t bytes, void** ptr)
{
	clear_error();
	if (!ptr) {
		set_error("rpwa_dev_alloc: null output");
		return RPWA_ERR_INVALID;
	}
	*ptr = nullptr;
	RPWA_HIP_CHECK(hipSetDevice(device));
	if (hipMalloc(ptr, bytes ? bytes : 1) != hipSuccess) {
		*ptr = nullptr;
		set_error("rpwa_dev_alloc: hipMalloc of " + std::to_string(bytes) + " bytes failed");
		return RPWA_ERR_NOMEM;
	}
	return RPWA_OK;
}

extern "C" int rpwa_dev_free(void* ptr)
{
	clear_error();
	RPWA_HIP_CHECK(hipFree(ptr));
	return RPWA_OK;
}

extern "C" int rpwa_dev_copy(void* dst, const void* src, uint64_t bytes, int kind)
{
	clear_error();
	if ((!dst || !src) && bytes) {
		set_error("rpwa_dev_copy: null pointer");
		return RPWA_ERR_INVALID;
	}
	hipMemcpyKind k;
	switch (kind) {
	case RPWA_COPY_H2D: k = hipMemcpyHostToDevice; break;
	case RPWA_COPY_D2H: k = hipMemcpyDeviceToHost; break;
	case RPWA_COPY_D2D: k = hipMemcpyDeviceToDevice; break;
	default:
		set_error("rpwa_dev_copy: unknown copy kind");
		return RPWA_ERR_INVALID;
	}
	RPWA_HIP_CHECK(hipMemcpy(dst, src, bytes, k));
	return RPWA_OK;
}

extern "C" int rpwa_dev_synchronize(int device)
{
	clear_error();
	RPWA_HIP_CHECK(hipSetDevice(device));
	RPWA_HIP_CHECK(hipDeviceSynchronize());
	return RPWA_OK;
}

extern "C" int rpwa_wave_nonzero_device(int device, const double* d_amps, uint64_t nmbEvents, uint32_t nmbWaves,
                                        uint32_t wave, int* nonzero)
{
	clear_error();
	if (!d_amps || !nonzero || nmbEvents == 0 || wave >= nmbWaves) {
		set_error("rpwa_wave_nonzero_device: null pointer, no events or wave index outside the row");
		return RPWA_ERR_INVALID;
	}
	RPWA_GUARD_BEGIN
	RPWA_HIP_CHECK(hipSetDevice(device));
	int nCu = 0;
	RPWA_HIP_CHECK(hipDeviceGetAttribute(&nCu, hipDeviceAttributeMultiprocessorCount, device));
	unsigned int* d_flag = nullptr;
	RPWA_HIP_CHECK(hipMalloc(&d_flag, sizeof(unsigned int)));
	unsigned int flag = 0;
	hipError_t e = hipMemcpy(d_flag, &flag, sizeof(flag), hipMemcpyHostToDevice);
	if (e == hipSuccess)
		e = launch_wave_nonzero(d_amps, nmbEvents, nmbWaves, wave, d_flag, nCu, nullptr);
	if (e == hipSuccess)
		e = hipMemcpy(&flag, d_flag, sizeof(flag), hipMemcpyDeviceToHost);
	hipFree(d_flag);
	RPWA_HIP_CHECK(e);
	*nonzero = flag ? 1 : 0;
	return RPWA_OK;
	RPWA_GUARD_END
}

extern "C" int rpwa_dev_info(int device, char* name, int nameLen, int* nCu, uint64_t* totalMem, int* clockKhz)
{
	clear_error();
	hipDeviceProp_t p;
	RPWA_HIP_CHECK(hipGetDeviceProperties(&p, device));
	if (name && nameLen > 0) {
		std::strncpy(name, p.gcnArchName, (size_t)nameLen - 1);
		name[nameLen - 1] = 0;
	}
	if (nCu)
		*nCu = p.multiProcessorCount;
	if (totalMem)
		*totalMem = p.totalGlobalMem;
	if (clockKhz)
		*clockKhz = p.clockRate;
	return RPWA_OK;
}

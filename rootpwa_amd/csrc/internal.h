// internal.h -- shared declarations of the gfx950 backend (not installed).
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <string>
#include <vector>

#include "rootpwa_amd.h"

namespace rpwa_amd {

// error reporting: every C entry point catches and converts to a status code
void set_error(const std::string& msg);
void clear_error();

#define RPWA_HIP_CHECK(expr)                                                            \
	do {                                                                                \
		hipError_t e_ = (expr);                                                         \
		if (e_ != hipSuccess) {                                                         \
			::rpwa_amd::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));   \
			return RPWA_ERR_HIP;                                                        \
		}                                                                               \
	} while (0)

// Highest rank held in registers by the event kernels; higher ranks run through
// the LDS-staged kernel in blocks of kMaxRank ranks (one context slot for all of them).
constexpr int kMaxRank = 4;
constexpr int kHighRankSlot = kMaxRank + 1;
inline int rank_slot(uint32_t rank) { return rank <= (uint32_t)kMaxRank ? (int)rank : kHighRankSlot; }

// launch modes of the LDS-staged kernel
constexpr int STAGE_FUSED = 0;  // all ranks in one launch: log sum, weight sum, gradient
constexpr int STAGE_SIGMA = 1;  // one rank block: add its sum_r |S_r|^2 to the per-event sigma
constexpr int STAGE_GRAD = 2;   // one rank block: gradient entries from the complete sigma

// Device layout of a likelihood context's amplitudes: one row per event; the
// row holds the refl -1 waves, then the refl +1 waves.  In mixed precision
// (complex64) each sector is padded to an even number of waves so that a lane
// can load two complex64 values of one sector with one 16-byte load.
struct AmpLayout {
	uint64_t nEvents = 0;
	uint32_t nW[2] = {0, 0};
	uint32_t wmax = 0;
	int precision = RPWA_PREC_FP64;
	uint32_t perSlot = 1;      // complex values per lane slot (1: fp64, 2: complex64)
	uint32_t secOff[2] = {0, 0};  // offset of each sector in a row [complex units]
	uint32_t rowLen = 0;       // row length [complex units]
	size_t bytes() const { return (size_t)nEvents * rowLen * (precision == RPWA_PREC_FP64 ? 16 : 8); }
};

// Per-launch description of the event tiling: a group of `lanes` lanes owns
// one event; lanes0 of them hold the refl -1 waves, lanes1 the refl +1 waves,
// each lane `slots` slots of perSlot complex values.
struct TileShape {
	int lanes = 16;
	int slots = 1;
	int lanes0 = 0;
	int lanes1 = 0;
	int events = 0;  // > 0: the LDS-staged kernel (one lane per event) with tiles of this many events
	int gm = 0;      // staged kernel: 1 = the gradient phase on FP32 MFMA (complex64 amplitudes)
};

// Geometry of the LDS-staged kernel (stage_kernel.hip) for one layout, tile size and rank.
struct StagePlan {
	int events = 0;
	int gm = 0;              // gradient phase on FP32 MFMA
	int threads = 256;       // workgroup size
	uint32_t sliceUnits = 0; // units per wave in the S phase
	int ranks = 1;           // ranks held by the kernel instantiation (1, 2 or 4)
	uint32_t rowUnits = 0;   // 16-byte units per amplitude row
	uint32_t rowUnitsP = 0;  // LDS row stride (odd)
	uint32_t gGroups0 = 0;   // unit groups of the refl -1 sector in the gradient phase
	uint32_t gGroups = 0;    // unit groups of both sectors
	uint32_t gSubs = 0;      // event subsets per unit group in the gradient phase
	uint32_t stageBytes = 0;
	size_t ldsBytes = 0;
};

struct EventArgs {
	const void* amps;
	uint64_t nEvents;
	uint32_t rowLen;
	uint32_t secOff0, secOff1;
	uint32_t nW0, nW1;
	uint32_t wmax;
	int rank;
	int lanes0, lanes1;
	const double2* prod;   // device [rank][2][wmax] complex
	const double2* tpad;   // device [rowLen][ranks of the staged kernel]: T per row column, 0 on padding
	const double2* tpadHost;  // the same on the host (the staged kernel takes it in its arguments when it fits)
	double flat2;
	double* partials;      // [nWg][nElem]
	uint32_t nElem;
	int stageMode = STAGE_FUSED;  // the staged kernel's launch mode
	uint32_t rank0 = 0;           // first rank of the block (STAGE_GRAD)
	uint32_t accumulate = 0;      // STAGE_SIGMA: add to sigma
	double* sigma = nullptr;      // per-event sigma of the rank blocks [nEvents]
};

// Number of doubles in a workgroup partial: lnL sum, lnL compensation,
// sum of weights -2/sigma, then [rank][2][wmax] complex derivatives.
inline uint32_t partial_elems(int rank, uint32_t wmax) { return 3u + 2u * (uint32_t)rank * 2u * wmax; }

// launchers (lik_kernels.hip).  A non-null start / stop event is bound to the
// launch itself (hipExtLaunchKernelGGL): it takes the kernel's own start / end
// time on the device, not the time the host queued it.
hipError_t launch_event_kernel(const AmpLayout& lay, const TileShape& ts, const EventArgs& a,
                               bool grad, int nWg, hipStream_t s, hipEvent_t start = nullptr,
                               hipEvent_t stop = nullptr);
int event_kernel_occupancy(const AmpLayout& lay, const TileShape& ts, int rank, int* wgPerCu);
bool stage_plan(const AmpLayout& lay, int events, int rank, StagePlan& p, int gm = 0);
// true if T of every row column fits into the staged kernel's arguments (no upload needed)
bool stage_t_inline(const AmpLayout& lay, int rank);
int stage_rank_cap(int rank);
hipError_t launch_stage_kernel(const AmpLayout& lay, int events, int gm, const EventArgs& a, int nWg, hipStream_t s,
                               hipEvent_t start = nullptr, hipEvent_t stop = nullptr);
int stage_kernel_occupancy(const AmpLayout& lay, int events, int gm, int rank, int* wgPerCu);
// doneFlag (one slot per block of the reduction, reduce_blocks(nElem)) may be null (no completion flags)
uint32_t reduce_blocks(uint32_t nElem);
hipError_t launch_event_reduce(const double* partials, int nWg, uint32_t nElem, double flat,
                               double* out, hipStream_t s,
                               unsigned long long* doneFlag = nullptr, unsigned long long seq = 0,
                               hipEvent_t start = nullptr, hipEvent_t stop = nullptr);
// weighted real rank-k update (integral.hip): full D x D matrix of sum_e wt_e x_e x_e^T
int syrk_run(int device, const double* d_x, uint64_t M, uint32_t D, const double* d_w, bool directWeights,
             const double* h_x, const double* h_w, std::vector<double>& full, double* wtot);
// *flag |= 1 if column `wave` of amps [nEvents][W] (complex128) has a non-zero element
hipError_t launch_wave_nonzero(const double* d_amps, uint64_t nEvents, uint32_t W, uint32_t wave,
                               unsigned int* flag, int nCu, hipStream_t s);
hipError_t launch_transform_amps(const AmpLayout& lay, const double* d_src, int srcLayout,
                                 const double* d_norm, void* d_dst, hipStream_t s);

}  // namespace rpwa_amd

struct rpwa_lik_ctx {
	std::atomic<int> refs{1};
	int device = 0;
	int precision = RPWA_PREC_FP64;
	hipStream_t stream = nullptr;
	// kernel timing: ev0 / evMid bracket the event kernel(s), evRed / ev1 the reduction
	hipEvent_t ev0 = nullptr, evMid = nullptr, evRed = nullptr, ev1 = nullptr;
	rpwa_amd::AmpLayout lay;
	// tile and persistent grid per rank (index = rank); fixed before the first
	// evaluation of that rank, so every result of a rank comes from one tile
	// (slot rank_slot(rank): ranks above kMaxRank share the last slot)
	rpwa_amd::TileShape tile[rpwa_amd::kHighRankSlot + 1];
	int nWg[rpwa_amd::kHighRankSlot + 1] = {0, 0, 0, 0, 0, 0};
	bool tileFixed[rpwa_amd::kHighRankSlot + 1] = {false, false, false, false, false, false};
	bool evaluated[rpwa_amd::kHighRankSlot + 1] = {false, false, false, false, false, false};
	bool autotune = true;
	bool loaded = false;
	int nCu = 0;
	void* d_amps = nullptr;
	uint32_t rankCap = 0;          // highest rank the buffers below are sized for
	double* d_sigma = nullptr;     // [nEvents] per-event sigma of the rank blocks (ranks above kMaxRank)
	double* d_prod = nullptr;      // [rankCap][2][wmax] complex, then the staged kernel's T per row column
	double* d_partials = nullptr;  // [nWg][partial_elems]
	double* d_out = nullptr;       // device view of h_out: the reduce kernel writes the results there
	double* h_prod = nullptr;      // pinned staging
	double* h_out = nullptr;       // reduced results (pinned, coherent, mapped into the device), then the flag
	unsigned long long* h_flag = nullptr;  // completion flags of the reduction's blocks after the results (host view) ...
	unsigned long long* d_flag = nullptr;  // ... and their device view
	uint32_t pendingBlocks = 0;            // blocks of the pending evaluation's reduction
	unsigned long long seq = 0;            // evaluations submitted
	uint32_t timing = 0;                   // kernel timing with events: every timing-th evaluation, 0 = off (rpwa_lik_set_timing)
	bool pendingTimed = false;             // the pending evaluation's launches carry the timing events
	size_t partialCap = 0;
	// pending asynchronous evaluation (rpwa_lik_submit / rpwa_lik_collect)
	bool pending = false;
	uint32_t pendingRank = 0;
	uint32_t pendingProd = 0;
	double kernelTime = 0;
	double eventTime = 0;
	double reduceTime = 0;
	uint64_t nEvals = 0;
};
